"""Host-side mirror of the reference's collision-operator interface over the C ABI.

CollisionOperator follows the reference's call sequence in the swarm runtime: a geometry epoch
(col_site_init + Morton build + col_cache_get + col_space_bin, src/swarm/swarm_runtime.cu:384-448)
followed by operator calls evolve_local_collisions(duration) (inc/swarm/_col_chain.cuh:108-539),
with the controller reset at output boundaries (swarm_runtime.cu:491-496). Errors raise GdxError
carrying the status code of the reference failure path. torch supplies device memory only.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def model_info(model: str) -> dict:
    lib = _lib.load(model)
    info = _lib.ModelInfo()
    _lib.check(lib, None, lib.gdx_col_model_info(C.byref(info)))
    out = {f: getattr(info, f) for f, _ in info._fields_}
    out["model"] = info.model.decode()
    return out


def _ptr(t) -> int:
    return t.data_ptr()


def _np_ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class CollisionOperator:
    """One swarm's collision operator on the current HIP device (one library per model directory)."""

    def __init__(self, model: str, n: int, total_dust_mass: float, seed: int = 0, defer_writeback: bool = False,
                 profile: bool = False, search: str = "morton", diagnostics: bool = False):
        self.model = model
        self.lib = _lib.load(model)
        self.info = model_info(model)
        self.n = int(n)
        flags = ((_lib.DEFER_WRITEBACK if defer_writeback else 0) | (_lib.PROFILE if profile else 0)
                 | (_lib.DIAGNOSTICS if diagnostics else 0))
        modes = {"morton": _lib.SEARCH_MORTON, "kdtree": _lib.SEARCH_KDTREE}
        if search not in modes:
            raise ValueError(f"search must be one of {sorted(modes)}")
        cfg = _lib.Config(self.n, float(total_dust_mass), int(seed) & (2**64 - 1), flags, modes[search])
        h = C.c_void_p()
        _lib.check(self.lib, None, self.lib.gdx_col_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = ()

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.gdx_col_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, fn, *args):
        _lib.check(self.lib, self.h, fn(self.h, *args))

    # ---- binding ----
    def bind_soa(self, x, y, z, size, numr) -> None:
        """Bind particle-order float64 CUDA tensors (azimuth, spherical radius, polar angle, size, number)."""
        for t in (x, y, z, size, numr):
            if not t.is_cuda or str(t.dtype) != "torch.float64" or not t.is_contiguous() or t.numel() != self.n:
                raise ValueError("state must be contiguous float64 device tensors of length n")
        self._keep = (x, y, z, size, numr)
        self._call(self.lib.gdx_col_bind_soa, *(_ptr(t) for t in (x, y, z, size, numr)))

    def bind_swarm(self, records) -> None:
        """Bind the reference's 64-byte swarm records (float64 device tensor of shape (n, 8))."""
        if tuple(records.shape) != (self.n, 8) or str(records.dtype) != "torch.float64" or not records.is_contiguous():
            raise ValueError("swarm records must be a contiguous float64 tensor of shape (n, 8)")
        self._keep = (records,)
        self._call(self.lib.gdx_col_bind_swarm, _ptr(records))

    # ---- operator ----
    def geometry(self) -> None:
        self._call(self.lib.gdx_col_geometry)

    def check_state(self) -> None:
        self._call(self.lib.gdx_col_check_state)

    def load_state(self) -> None:
        self._call(self.lib.gdx_col_load_state)

    def store_state(self) -> None:
        self._call(self.lib.gdx_col_store_state)

    def evolve(self, duration: float) -> dict:
        st = _lib.Stats()
        self._call(self.lib.gdx_col_operator, C.c_double(duration), C.byref(st))
        return st.as_dict()

    def reset_controller(self) -> None:
        self._call(self.lib.gdx_col_reset_controller)

    def synchronize(self) -> None:
        self._call(self.lib.gdx_col_synchronize)

    # ---- diagnostics (diagnostics=True) ----
    def open_log(self, directory: str) -> None:
        """Write one schema-1 JSON line per operator call to <directory>collision_local_<stamp>.jsonl."""
        self._call(self.lib.gdx_col_open_log, directory.encode())

    def set_clock(self, clock_sim: float) -> None:
        self._call(self.lib.gdx_col_set_clock, float(clock_sim))

    def save_controller(self, path: str) -> None:
        """Controller archive (schema 2) of every bath since the last reset_controller()."""
        self._call(self.lib.gdx_col_save_controller, str(path).encode())

    def event_totals(self) -> tuple[np.ndarray, np.ndarray]:
        """Event counts and summed log-mass change per outcome category of the last operator call."""
        counts = np.zeros(7, dtype=np.uint64)
        log_mass = np.zeros(7, dtype=np.float64)
        self._call(self.lib.gdx_col_event_totals, _np_ptr(counts), _np_ptr(log_mass))
        return counts, log_mass

    def rate_pass(self) -> None:
        self._call(self.lib.gdx_col_rate_pass)

    def kernel_times(self) -> dict:
        """Accumulated HIP-event time (ms), launches and owners per operator kernel (profile=True)."""
        nk = len(_lib.KERNELS)
        ms = np.zeros(nk, np.float64)
        la = np.zeros(nk, np.int64)
        ow = np.zeros(nk, np.int64)
        self._call(self.lib.gdx_col_kernel_times, _np_ptr(ms), _np_ptr(la), _np_ptr(ow))
        return {k: {"ms": float(ms[i]), "launches": int(la[i]), "owners": int(ow[i])}
                for i, k in enumerate(_lib.KERNELS)}

    def reset_kernel_times(self) -> None:
        self._call(self.lib.gdx_col_reset_kernel_times)

    # ---- inspection ----
    def active_count(self) -> int:
        v = C.c_int()
        self._call(self.lib.gdx_col_active_count, C.byref(v))
        return v.value

    def neighbors(self) -> tuple[np.ndarray, np.ndarray]:
        k = self.info["n_k"]
        codes = np.empty((self.n, k), dtype=np.int32)
        meas = np.empty(self.n, dtype=np.float64)
        self._call(self.lib.gdx_col_get_neighbors, _np_ptr(codes), _np_ptr(meas))
        return codes, meas

    def sites(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        pts = np.empty((self.n, 4), dtype=np.float32)
        act = np.empty(self.n, dtype=np.uint8)
        grp = np.empty(self.n, dtype=np.uint8)
        self._call(self.lib.gdx_col_get_sites, _np_ptr(pts), _np_ptr(act), _np_ptr(grp))
        return pts, act, grp

    def morton(self) -> tuple[np.ndarray, np.ndarray]:
        keys = np.empty(self.n, dtype=np.uint64)
        idx = np.empty(self.n, dtype=np.int32)
        self._call(self.lib.gdx_col_get_morton, _np_ptr(keys), _np_ptr(idx))
        return keys, idx

    def kdtree(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """k-d tree of the last geometry epoch: particle index and split dimension per node, world box."""
        idx = np.empty(self.n, dtype=np.int32)
        dim = np.empty(self.n, dtype=np.int32)
        world = np.empty(6, dtype=np.float32)
        self._call(self.lib.gdx_col_get_kdtree, _np_ptr(idx), _np_ptr(dim), _np_ptr(world))
        return idx, dim, world

    def cached(self) -> np.ndarray:
        out = np.empty((self.n, 4), dtype=np.float64)
        self._call(self.lib.gdx_col_get_cached, _np_ptr(out))
        return out

    def rng(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint64)
        self._call(self.lib.gdx_col_get_rng, _np_ptr(out))
        return out

    def set_rng(self, counters: np.ndarray) -> None:
        a = np.ascontiguousarray(counters, dtype=np.uint64)
        self._call(self.lib.gdx_col_set_rng, _np_ptr(a))

    def graph(self) -> np.ndarray:
        out = np.empty(self.info["local_groups"], dtype=np.uint64)
        self._call(self.lib.gdx_col_get_graph, _np_ptr(out))
        return out


def evolve_batch(ops: list, duration: float) -> list[dict]:
    """One operator call on each of several CollisionOperator instances of the same model (independent
    swarms), sharing every kernel launch; returns one stats dict per operator."""
    if not ops:
        return []
    lib = ops[0].lib
    if any(op.lib is not lib for op in ops):
        raise ValueError("all operators of a batch must use the same model library")
    handles = (C.c_void_p*len(ops))(*[op.h for op in ops])
    stats = (_lib.Stats*len(ops))()
    _lib.check(lib, ops[0].h, lib.gdx_col_operator_batch(handles, len(ops), float(duration), stats))
    return [s.as_dict() for s in stats]


class MortonIndex:
    """Stand-alone adaptive Morton index (morton_index::build, inc/swarm/morton/morton_index.cuh:168-229)."""

    def __init__(self, model: str, points4, active=None, origin=(0.0, 0.0, 0.0), width=1.0, dim=3,
                 leaf_target=128, max_level=20):
        self.lib = _lib.load(model)
        self.n = int(points4.shape[0])
        self._pts = points4
        self._act = active
        h = C.c_void_p()
        _lib.check(self.lib, None, self.lib.gdx_morton_build(
            _ptr(points4), _ptr(active) if active is not None else None, self.n, C.c_float(origin[0]),
            C.c_float(origin[1]), C.c_float(origin[2]), C.c_float(width), dim, leaf_target, max_level, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.gdx_morton_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self) -> dict:
        nn = C.c_int()
        _lib.check(self.lib, None, self.lib.gdx_morton_node_count(self.h, C.byref(nn)))
        m = nn.value
        out = {
            "keys": np.empty(self.n, np.uint64), "idx": np.empty(self.n, np.int32),
            "lower": np.empty((m, 3), np.float32), "width": np.empty(m, np.float32), "begin": np.empty(m, np.int32),
            "count": np.empty(m, np.int32), "child": np.empty((m, 8), np.int32), "child_count": np.empty(m, np.int32),
        }
        _lib.check(self.lib, None, self.lib.gdx_morton_get(self.h, *(_np_ptr(out[k]) for k in (
            "keys", "idx", "lower", "width", "begin", "count", "child", "child_count"))))
        return out

    def knn(self, queries4, k: int):
        import torch
        nq = int(queries4.shape[0])
        codes = torch.empty((nq, k), dtype=torch.int32, device=queries4.device)
        d2 = torch.empty((nq, k), dtype=torch.float32, device=queries4.device)
        _lib.check(self.lib, None, self.lib.gdx_morton_knn(self.h, _ptr(queries4), nq, k, _ptr(codes), _ptr(d2)))
        return codes.cpu().numpy(), d2.cpu().numpy()
