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

    def set_stream(self, stream) -> None:
        """Launch on a HIP stream (a torch.cuda.Stream or a raw handle; None: the null stream)."""
        handle = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        self._call(self.lib.gdx_col_set_stream, C.c_void_p(handle))

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

    def record_info(self) -> dict:
        """Indexed records of the last Morton epoch (n plus periodic ghosts in a wedge), whether images of
        one particle are disjoint, and the largest search cap."""
        r, u, q = C.c_int(0), C.c_int(0), C.c_float(0.0)
        self._call(self.lib.gdx_col_record_info, C.byref(r), C.byref(u), C.byref(q))
        return {"records": r.value, "unique_ids": bool(u.value), "q_max": q.value}

    def morton(self) -> tuple[np.ndarray, np.ndarray]:
        m = self.record_info()["records"]
        keys = np.empty(m, dtype=np.uint64)
        idx = np.empty(m, dtype=np.int32)
        self._call(self.lib.gdx_col_get_morton, _np_ptr(keys), _np_ptr(idx))
        return keys, idx

    def kdtree(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """k-d tree of the last geometry epoch: particle index and split dimension per node, world box."""
        m = self.record_info()["records"]
        idx = np.empty(m, dtype=np.int32)
        dim = np.empty(m, dtype=np.int32)
        world = np.empty(6, dtype=np.float32)
        self._call(self.lib.gdx_col_get_kdtree, _np_ptr(idx), _np_ptr(dim), _np_ptr(world))
        return idx, dim, world

    def cached(self) -> np.ndarray:
        out = np.empty((self.n, 4), dtype=np.float64)
        self._call(self.lib.gdx_col_get_cached, _np_ptr(out))
        return out

    def pair_rates(self) -> np.ndarray:
        """n x N_K bath-start pair rates (particle order, slot order); publishes the reservoir first."""
        out = np.empty((self.n, self.info["n_k"]), dtype=np.float64)
        self._call(self.lib.gdx_col_pair_rates, _np_ptr(out))
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

    # ---- search tuning ----
    def set_leaf_target(self, leaf_target: int) -> None:
        self._call(self.lib.gdx_col_set_leaf_target, int(leaf_target))

    def set_profile(self, on: bool) -> None:
        """Kernel timing (HIP events per launch) of later operator calls on or off."""
        self._call(self.lib.gdx_col_set_profile, int(bool(on)))

    def set_bound_reuse(self, on: bool) -> None:
        """Learned: keep the bounds of owners whose inputs did not change (default on); results do not
        depend on it."""
        self._call(self.lib.gdx_col_set_bound_reuse, int(bool(on)))

    def set_fixed_bath(self, duration: float) -> None:
        """Test access: every bath of later calls lasts min(duration, horizon); 0 restores the controller."""
        self._call(self.lib.gdx_col_set_fixed_bath, float(duration))

    def set_exact_walk(self, on: bool) -> None:
        """Test access: the chain picks every partner from the walk's sequential partial sums (default
        off: tree-summed partials decide wherever provably on the same side of the target). Results do
        not depend on it."""
        self._call(self.lib.gdx_col_set_exact_walk, int(bool(on)))

    def set_narrow_waves(self, owners: int) -> None:
        """Refresh waves of at most `owners` due owners take the narrow launch shapes (four lanes per
        plain owner in the rate pass, one owner per lane in the screen); 0 always takes the wide ones.
        Results do not depend on it."""
        self._call(self.lib.gdx_col_set_narrow_waves, int(owners))

    def autotune(self, candidates=(16, 32, 64, 128, 256), repeats: int = 2) -> dict:
        """Time the geometry epoch per Morton leaf target, keep the fastest (a valid epoch is left built
        with it); returns {"best": leaf, "ms": {leaf: ms}}."""
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c), np.float64)
        best = C.c_int()
        self._call(self.lib.gdx_col_autotune, _np_ptr(c), len(c), int(repeats), C.byref(best), _np_ptr(ms))
        return {"best": best.value, "ms": {int(k): float(v) for k, v in zip(c, ms)}}

    def set_topk_tile(self, tile: int) -> None:
        """Top-K selection of the Morton search: 0 register insertion, 8/16/32 LDS-tiled merges."""
        self._call(self.lib.gdx_col_set_topk_tile, int(tile))

    def autotune_tile(self, candidates=(0, 8, 16, 32), repeats: int = 2) -> dict:
        """Time the geometry epoch per top-K tile at the current leaf target, keep the fastest."""
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c), np.float64)
        best = C.c_int()
        self._call(self.lib.gdx_col_autotune_tile, _np_ptr(c), len(c), int(repeats), C.byref(best), _np_ptr(ms))
        return {"best": best.value, "ms": {int(k): float(v) for k, v in zip(c, ms)}}

    def set_topk_block(self, waves: int) -> None:
        """Block shape of the Morton search: 1, 2 or 4 wavefronts per block of the top-K kernel."""
        self._call(self.lib.gdx_col_set_topk_block, int(waves))

    def autotune_block(self, candidates=(1, 2, 4), repeats: int = 2) -> dict:
        """Time the geometry epoch per top-K block shape at the current leaf target and tile, keep the
        fastest."""
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c), np.float64)
        best = C.c_int()
        self._call(self.lib.gdx_col_autotune_block, _np_ptr(c), len(c), int(repeats), C.byref(best), _np_ptr(ms))
        return {"best": best.value, "ms": {int(k): float(v) for k, v in zip(c, ms)}}

    # ---- learned majorant ----
    def majorant_sample(self):
        """Features (n x 8, float32), partner scale S and exact bath-start rate R per particle at the
        current state (zero rows for inactive particles), as device tensors."""
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        feats = torch.empty((self.n, _lib.MAJORANT_FEATURES), dtype=torch.float32, device=dev)
        scale = torch.empty(self.n, dtype=torch.float64, device=dev)
        rate = torch.empty(self.n, dtype=torch.float64, device=dev)
        self._call(self.lib.gdx_col_majorant_sample, _ptr(feats), _ptr(scale), _ptr(rate))
        return feats, scale, rate

    def use_majorant(self, majorant, safety: float = 1.2, min_owners: int = 0) -> None:
        """Draw event times by thinning against the majorant (None: the plain method). Swarms with fewer
        than min_owners active owners (0: the library default, 65,536) keep the plain method."""
        self._call(self.lib.gdx_col_use_majorant, majorant.h if majorant is not None else None, float(safety),
                   int(min_owners))
        self._majorant = majorant


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


def evolve_ensemble(ops: list, duration: float, steps: int = 1, streams: int = 4) -> list[list[dict]]:
    """`steps` operator calls on every operator of an ensemble of independent realizations, split over
    `streams` concurrent groups (one HIP stream and host thread each, advancing independently); the same
    results as separate calls. Returns the stats per step, one dict per operator in `ops` order."""
    if not ops:
        return []
    lib = ops[0].lib
    if any(op.lib is not lib for op in ops):
        raise ValueError("all operators of an ensemble must use the same model library")
    n = len(ops)
    handles = (C.c_void_p*n)(*[op.h for op in ops])
    stats = (_lib.Stats*(n*steps))()
    _lib.check(lib, ops[0].h, lib.gdx_col_operator_ensemble(handles, n, float(duration), int(steps), int(streams),
                                                           stats))
    return [[stats[s*n + i].as_dict() for i in range(n)] for s in range(steps)]


class Majorant:
    """Learned rate majorant (no reference counterpart): an 8-64-64-1 BF16 perceptron trained and run by
    the library's MFMA kernels. Feature rows and targets are device tensors (torch supplies memory)."""

    def __init__(self, model: str, seed: int = 0):
        self.model = model
        self.lib = _lib.load(model)
        h = C.c_void_p()
        rc = self.lib.gdx_col_majorant_create(int(seed) & (2**64 - 1), C.byref(h))
        self._check(rc, None)
        self.h = h

    def _check(self, rc: int, h) -> None:
        if rc != 0:
            msg = self.lib.gdx_col_majorant_last_error(h)
            raise _lib.GdxError(rc, msg.decode() if msg else "")

    def _call(self, fn, *args):
        self._check(fn(self.h, *args), self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.gdx_col_majorant_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> tuple[np.ndarray, np.ndarray]:
        p = np.empty(_lib.MAJORANT_PARAMS, np.float32)
        norm = np.empty(2*_lib.MAJORANT_FEATURES, np.float32)
        self._call(self.lib.gdx_col_majorant_get_params, _np_ptr(p), _np_ptr(norm))
        return p, norm

    def set_params(self, params=None, norm=None) -> None:
        pa = None if params is None else np.ascontiguousarray(params, np.float32)
        na = None if norm is None else np.ascontiguousarray(norm, np.float32)
        self._call(self.lib.gdx_col_majorant_set_params, None if pa is None else _np_ptr(pa),
                   None if na is None else _np_ptr(na))

    def fit_normalization(self, features) -> np.ndarray:
        """Set the input normalization to the per-feature mean and standard deviation (1 where constant)."""
        x = features.detach().double().cpu().numpy() if hasattr(features, "detach") else np.asarray(features)
        mu = x.mean(axis=0)
        sd = x.std(axis=0)
        sd = np.where(sd > 1e-12, sd, 1.0)
        norm = np.concatenate([mu, sd]).astype(np.float32)
        self.set_params(None, norm)
        return norm

    def fit_output(self, log_target) -> float:
        """Start the output bias at the mean log target, softplus(b3) = exp(mean), so that training
        begins at the right scale; returns b3."""
        import math
        m = float(log_target.double().mean()) if hasattr(log_target, "double") else float(np.mean(log_target))
        b3 = m if m < -20.0 else (math.log(math.expm1(math.exp(m))) if m < 20.0 else math.exp(m))
        p, _ = self.params()
        p[-1] = b3
        self.set_params(p, None)
        return b3

    def calibrate(self, features, log_target, quantile: float = 0.999, floor: float = 1.2) -> float:
        """Safety factor for the bound: the `quantile` of exact / predicted over the samples, at least
        `floor` (the factor the bound is used with in gdx_col_use_majorant)."""
        import torch
        y = self.infer(features).double()
        r = torch.exp(log_target.double())/y
        r = r[torch.isfinite(r)]
        if r.numel() == 0:
            return floor
        q = float(torch.sort(r).values[min(r.numel() - 1, int(quantile*r.numel()))])
        return max(floor, q)

    def fit(self, features, log_target, epochs: int = 20, batch: int = 8192, lr: float = 1e-3, seed: int = 0,
            quantile: float = 0.999, floor: float = 1.2) -> tuple[np.ndarray, float]:
        """Normalization, output bias, Adam training and the calibrated safety factor in one call."""
        self.fit_normalization(features)
        self.fit_output(log_target)
        losses = self.train(features, log_target, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return losses, self.calibrate(features, log_target, quantile, floor)

    def infer(self, features):
        import torch
        n = features.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=features.device)
        self._call(self.lib.gdx_col_majorant_infer, _ptr(features.contiguous()), n, _ptr(out))
        return out

    def grad(self, features, log_target) -> tuple[np.ndarray, float]:
        g = np.empty(_lib.MAJORANT_PARAMS, np.float32)
        loss = C.c_double()
        self._call(self.lib.gdx_col_majorant_grad, _ptr(features.contiguous()), _ptr(log_target.contiguous()),
                   int(features.shape[0]), _np_ptr(g), C.byref(loss))
        return g, loss.value

    def train(self, features, log_target, epochs: int = 20, batch: int = 8192, lr: float = 1e-3,
              seed: int = 0) -> np.ndarray:
        cfg = _lib.TrainConfig(int(epochs), int(batch), float(lr), 0.0, 0.0, 0.0, int(seed) & (2**64 - 1))
        losses = np.zeros(max(1, epochs), np.float32)
        self._call(self.lib.gdx_col_majorant_train, _ptr(features.contiguous()), _ptr(log_target.contiguous()),
                   int(features.shape[0]), C.byref(cfg), _np_ptr(losses))
        return losses[:epochs]


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
