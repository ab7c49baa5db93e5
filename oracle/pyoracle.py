"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
P = C.c_void_p


class ModelInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("n_k", "dim", "coag_kernel", "n_x", "n_y", "n_z", "local_groups", "col_bin_s",
                                       "raw_bins", "morton_leaf_target", "morton_max_level", "x_wedge")] + \
               [(n, C.c_double) for n in ("y_min", "y_max", "z_min", "z_max", "x_min", "x_max", "h_search", "v_frag",
                                          "col_bath_max", "col_bath_eps", "col_bath_alpha", "init_smin", "rho_0")]


class Stats(C.Structure):
    _fields_ = [("waves", C.c_int), ("launches_equiv", C.c_int), ("errors", C.c_int), ("events", C.c_longlong),
                ("owner_updates", C.c_longlong), ("pair_evals", C.c_longlong), ("baths", C.c_longlong),
                ("dt_col", C.c_double), ("clock_dyn", C.c_double)]


def _p(a: np.ndarray):
    return a.ctypes.data_as(P)


class Oracle:
    def __init__(self, model: str, path: Path | None = None):
        path = path or HERE / "lib" / f"liboracle_{model}.so"
        if not path.exists():
            raise RuntimeError(f"oracle for {model} is not built: {path} (python oracle/build.py)")
        self.lib = L = C.CDLL(str(path))
        self.model = model
        d, i, u64 = C.c_double, C.c_int, C.c_uint64
        L.orc_uniform.restype = d
        L.orc_uniform.argtypes = [C.c_uint32, u64, u64]
        for name in ("orc_vrel_t", "orc_vrel_pair", "orc_ball_measure"):
            getattr(L, name).restype = d
        L.orc_vrel_t.argtypes = [d]*5
        L.orc_vrel_pair.argtypes = [d]*4
        L.orc_ball_measure.argtypes = [d]*3
        L.orc_outcome_moments.restype = d
        L.orc_outcome_moments.argtypes = [d, d, i, C.POINTER(d), C.POINTER(d), C.POINTER(d)]
        L.orc_sample_outcome.restype = d
        L.orc_sample_outcome.argtypes = [d, d, i, d, C.POINTER(i), C.POINTER(d)]
        L.orc_chain_rate.restype = d
        L.orc_chain_rate.argtypes = [d]*8 + [C.POINTER(d)]
        L.orc_stokes.restype = d
        L.orc_stokes.argtypes = [d]*3
        L.orc_partner_walk.restype = i
        L.orc_partner_walk.argtypes = [P, i, d, i]
        L.orc_site_init.restype = i
        L.orc_site_init.argtypes = [i] + [P]*9
        L.orc_morton_keys.argtypes = [i, P, C.c_float, C.c_float, C.c_float, C.c_float, i, i, P]
        L.orc_sort_perm.argtypes = [i, P, P]
        L.orc_hierarchy_build.restype = i
        L.orc_hierarchy_build.argtypes = [i, P, C.c_float, C.c_float, C.c_float, C.c_float, i, i, i, i] + [P]*6
        L.orc_topk_brute.argtypes = [i, P, P, i, P, P, P, i, P, P, i]
        L.orc_image_dist_min.restype = C.c_float
        L.orc_image_dist_min.argtypes = []
        L.orc_wedge_records.restype = i
        L.orc_wedge_records.argtypes = [i, P, P, C.c_float, P, P]
        L.orc_topk_records.argtypes = [i, P, P, P, i, P, P, i, i, P, P, i]
        L.orc_topk_tree.argtypes = [i, P, P, P, i, P, P, P, P, P, P, i, i, i, P, i, P, P, P, i]
        L.orc_space_bin.argtypes = [i, P, P, P, P]
        L.orc_kd_build.argtypes = [i, P, P, P, P]
        L.orc_kd_query.argtypes = [i, P, P, P, P, P, i, P, i, P, P, P, i]
        L.orc_kd_query_wedge.argtypes = [i, P, P, P, P, P, i, P, i, P, P, P, i]
        L.orc_kd_images.argtypes = [i, P, P]
        L.orc_change_limit.restype = i
        L.orc_change_limit.argtypes = [d, d, d, d, C.POINTER(d), C.POINTER(i)]
        L.orc_build_binmap.restype = i
        L.orc_build_binmap.argtypes = [P, i, P]
        L.orc_choose_bath.restype = i
        L.orc_choose_bath.argtypes = [P, i, d, d, C.POINTER(d), C.POINTER(i)]
        L.orc_finish_bath.restype = i
        L.orc_finish_bath.argtypes = [P, i, P, P]
        L.orc_schedule_run.restype = i
        L.orc_schedule_run.argtypes = [i, d, P, P, i, P, P, P]
        L.orc_ctl_size.restype = i
        L.orc_ctl_init.argtypes = [P]
        L.orc_operator.restype = i
        L.orc_operator.argtypes = [i, P, P, P, P, P, P, P, P, P, u64, d, d, P, i, C.POINTER(Stats), C.c_char_p, i, P]
        L.orc_diag_new.restype = P
        L.orc_diag_free.argtypes = [P]
        L.orc_diag_reset.argtypes = [P]
        L.orc_diag_json.restype = i
        L.orc_diag_json.argtypes = [P, i, C.c_char_p, i]
        L.orc_bath_rates.argtypes = [i, P, P, P, P, P, P, P, d, P, i]
        L.orc_pair_rates.argtypes = [i, P, P, P, P, P, P, P, d, P, i]
        info = ModelInfo()
        L.orc_get_model_info(C.byref(info))
        self.info = {f: getattr(info, f) for f, _ in info._fields_}

    # ---- scalar physics ----
    def uniform(self, pidx, seed, counter):
        return self.lib.orc_uniform(pidx, seed, counter)

    def philox(self, ctr, key):
        c = (C.c_uint32*4)(*ctr)
        k = (C.c_uint32*2)(*key)
        o = (C.c_uint32*4)()
        self.lib.orc_philox4x32(c, k, o)
        return list(o)

    def vrel_t(self, R, si, sj, h, sig):
        return self.lib.orc_vrel_t(R, si, sj, h, sig)

    def vrel_pair(self, y, z, si, sj):
        return self.lib.orc_vrel_pair(y, z, si, sj)

    def ball_measure(self, y, z, radius):
        return self.lib.orc_ball_measure(y, z, radius)

    def outcome_moments(self, si, sj, high):
        m, s, x = C.c_double(), C.c_double(), C.c_double()
        f = self.lib.orc_outcome_moments(si, sj, int(high), C.byref(m), C.byref(s), C.byref(x))
        return f, m.value, s.value, x.value

    def sample_outcome(self, si, sj, high, u):
        cat, lm = C.c_int(), C.c_double()
        v = self.lib.orc_sample_outcome(si, sj, int(high), u, C.byref(cat), C.byref(lm))
        return v, cat.value, lm.value

    def chain_rate(self, yi, zi, yj, zj, si, sj, nj, lam):
        v = C.c_double()
        r = self.lib.orc_chain_rate(yi, zi, yj, zj, si, sj, nj, lam, C.byref(v))
        return r, v.value

    def stokes(self, y, z, size):
        return self.lib.orc_stokes(y, z, size)

    def partner_walk(self, rate, target, last_positive=-3):
        r = np.ascontiguousarray(rate, dtype=np.float64)
        return self.lib.orc_partner_walk(_p(r), len(r), float(target), int(last_positive))

    # ---- geometry ----
    def site_init(self, x, y, z, size, numr):
        n = len(x)
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, size, numr)]
        pts = np.empty((n, 3), np.float32)
        posx = np.empty(n, np.float32)
        q = np.empty(n, np.float32)
        act = np.empty(n, np.uint8)
        bad = self.lib.orc_site_init(n, *[_p(v) for v in a], _p(pts), _p(posx), _p(q), _p(act))
        return pts, posx, q, act, bad

    def morton_keys(self, pts, origin, width, max_level, dim):
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        keys = np.empty(len(pts), np.uint64)
        self.lib.orc_morton_keys(len(pts), _p(pts), *[C.c_float(v) for v in origin], C.c_float(width), max_level, dim,
                                 _p(keys))
        return keys

    def sort_perm(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        perm = np.empty(len(keys), np.int32)
        self.lib.orc_sort_perm(len(keys), _p(keys), _p(perm))
        return perm

    def hierarchy(self, sorted_keys, origin, width, dim, leaf_target, max_level, cap=None):
        sk = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
        n = len(sk)
        cap = cap or max(1024, 64*n + 64)
        out = {"lower": np.empty((cap, 3), np.float32), "width": np.empty(cap, np.float32),
               "begin": np.empty(cap, np.int32), "count": np.empty(cap, np.int32),
               "child": np.empty((cap, 8), np.int32), "child_count": np.empty(cap, np.int32)}
        m = self.lib.orc_hierarchy_build(n, _p(sk), *[C.c_float(v) for v in origin], C.c_float(width), dim,
                                         leaf_target, max_level, cap, *[_p(out[k]) for k in (
                                             "lower", "width", "begin", "count", "child", "child_count")])
        if m < 0:
            raise RuntimeError("hierarchy capacity exceeded")
        return {k: v[:m].copy() for k, v in out.items()}

    def topk_brute(self, pts, active, query_pts, query_q, k, threads=0):
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        act = np.ascontiguousarray(active, dtype=np.uint8)
        qp = np.ascontiguousarray(query_pts, dtype=np.float32)
        qq = np.ascontiguousarray(query_q, dtype=np.float32)
        nq = len(qp)
        codes = np.empty((nq, k), np.int32)
        d2 = np.empty((nq, k), np.float32)
        qi = np.zeros(nq, np.int32)
        self.lib.orc_topk_brute(len(pts), _p(pts), _p(act), nq, _p(qi), _p(qp), _p(qq), k, _p(codes), _p(d2), threads)
        return codes, d2

    def image_dist_min(self) -> np.float32:
        return np.float32(self.lib.orc_image_dist_min())

    def wedge_records(self, pts, posx, q_max):
        """Physical records then seam ghosts (float xyz) and their codes 3i + image."""
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        posx = np.ascontiguousarray(posx, dtype=np.float32)
        n = len(pts)
        rec = np.empty((3*n, 3), np.float32)
        code = np.empty(3*n, np.int32)
        m = self.lib.orc_wedge_records(n, _p(pts), _p(posx), C.c_float(q_max), _p(rec), _p(code))
        return rec[:m].copy(), code[:m].copy()

    def topk_records(self, rec, code, active, query_pts, query_q, k, dedup, threads=0):
        rec = np.ascontiguousarray(rec, dtype=np.float32)
        code = np.ascontiguousarray(code, dtype=np.int32)
        act = np.ascontiguousarray(active, dtype=np.uint8)
        qp = np.ascontiguousarray(query_pts, dtype=np.float32)
        qq = np.ascontiguousarray(query_q, dtype=np.float32)
        nq = len(qp)
        codes = np.empty((nq, k), np.int32)
        d2 = np.empty((nq, k), np.float32)
        self.lib.orc_topk_records(len(rec), _p(rec), _p(code), _p(act), nq, _p(qp), _p(qq), k, int(bool(dedup)),
                                  _p(codes), _p(d2), threads)
        return codes, d2

    def topk_tree(self, pts, active, sorted_idx, nodes, dim, max_level, query_idx, k, search_dist, threads=0):
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        act = np.ascontiguousarray(active, dtype=np.uint8)
        si = np.ascontiguousarray(sorted_idx, dtype=np.int32)
        qi = np.ascontiguousarray(query_idx, dtype=np.int32)
        sd = np.ascontiguousarray(search_dist, dtype=np.float32)
        nq = len(qi)
        codes = np.empty((nq, k), np.int32)
        d2 = np.empty((nq, k), np.float32)
        nd = {kk: np.ascontiguousarray(v) for kk, v in nodes.items()}
        self.lib.orc_topk_tree(len(pts), _p(pts), _p(act), _p(si), len(nd["width"]), _p(nd["lower"]), _p(nd["width"]),
                               _p(nd["begin"]), _p(nd["count"]), _p(nd["child"]), _p(nd["child_count"]), dim,
                               max_level, nq, _p(qi), k, _p(sd), _p(codes), _p(d2), threads)
        return codes, d2

    def kd_build(self, pts):
        """Left-balanced k-d tree over all points: (particle index per node, split dim per node, world box)."""
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        n = len(pts)
        idx = np.empty(n, np.int32)
        dim = np.empty(n, np.int32)
        world = np.empty(6, np.float32)
        self.lib.orc_kd_build(n, _p(pts), _p(idx), _p(dim), _p(world))
        return idx, dim, world

    def kd_query(self, pts, active, tree, query_idx, k, search_dist, threads=0):
        """Heap-ordered top-K rows (codes 3*j, -1 empty) and d^2 of the queries over a kd_build tree."""
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        act = np.ascontiguousarray(active, dtype=np.uint8)
        idx, dim, world = (np.ascontiguousarray(a) for a in tree)
        qi = np.ascontiguousarray(query_idx, dtype=np.int32)
        sd = np.ascontiguousarray(search_dist, dtype=np.float32)
        nq = len(qi)
        codes = np.empty((nq, k), np.int32)
        d2 = np.empty((nq, k), np.float32)
        self.lib.orc_kd_query(len(pts), _p(pts), _p(act), _p(idx), _p(dim), _p(world), nq, _p(qi), k, _p(sd),
                              _p(codes), _p(d2), threads)
        return codes, d2

    def kd_images(self, pts):
        """Physical points, then all rotated by -width (image 1), then by +width (image 2)."""
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        out = np.empty((3*len(pts), 3), np.float32)
        self.lib.orc_kd_images(len(pts), _p(pts), _p(out))
        return out

    def kd_query_wedge(self, pts3, active, tree, query_idx, k, search_dist, threads=0):
        """Heap-ordered rows (codes 3*j + image) over a kd_build tree of kd_images points; the
        queries, activity and search caps index physical particles."""
        pts3 = np.ascontiguousarray(pts3, dtype=np.float32)
        act = np.ascontiguousarray(active, dtype=np.uint8)
        idx, dim, world = (np.ascontiguousarray(a) for a in tree)
        qi = np.ascontiguousarray(query_idx, dtype=np.int32)
        sd = np.ascontiguousarray(search_dist, dtype=np.float32)
        nq = len(qi)
        codes = np.empty((nq, k), np.int32)
        d2 = np.empty((nq, k), np.float32)
        self.lib.orc_kd_query_wedge(len(pts3) // 3, _p(pts3), _p(act), _p(idx), _p(dim), _p(world), nq, _p(qi), k,
                                    _p(sd), _p(codes), _p(d2), threads)
        return codes, d2

    def space_bin(self, x, y, z):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
        out = np.empty(len(x), np.int32)
        self.lib.orc_space_bin(len(x), *[_p(v) for v in a], _p(out))
        return out

    # ---- controller ----
    def change_limit(self, A, B, eps, horizon):
        d, r = C.c_double(), C.c_int()
        rc = self.lib.orc_change_limit(A, B, eps, horizon, C.byref(d), C.byref(r))
        if rc:
            raise RuntimeError("change_limit rejected its inputs")
        return d.value, r.value

    def build_binmap(self, raw_count):
        rc = np.ascontiguousarray(raw_count, dtype=np.int32)
        out = np.empty(len(rc), np.int32)
        m = self.lib.orc_build_binmap(_p(rc), len(rc), _p(out))
        return m, out

    def choose_bath(self, bins6, remaining, limit_scale):
        b = np.ascontiguousarray(bins6, dtype=np.float64).reshape(-1, 6)
        d, r = C.c_double(), C.c_int()
        rc = self.lib.orc_choose_bath(_p(b), len(b), remaining, limit_scale, C.byref(d), C.byref(r))
        if rc:
            raise RuntimeError("choose_bath rejected its inputs")
        return d.value, r.value

    def finish_bath(self, audit15, state4):
        a = np.ascontiguousarray(audit15, dtype=np.float64).reshape(-1, 15)
        st = np.ascontiguousarray(state4, dtype=np.float64).copy()
        res = np.empty(10, np.float64)
        rc = self.lib.orc_finish_bath(_p(a), len(a), _p(st), _p(res))
        if rc:
            raise RuntimeError("finish_bath rejected its inputs")
        return res, st

    def schedule_run(self, horizon, requested, edges, max_waves=1 << 16):
        req = np.ascontiguousarray(requested, dtype=np.float64)
        e = np.ascontiguousarray(edges, dtype=np.uint32)
        n = len(req)
        ticks = np.empty(max_waves, np.uint64)
        due = np.empty(max_waves, np.int32)
        levels = np.empty(n, np.int32)
        w = self.lib.orc_schedule_run(n, horizon, _p(req), _p(e), max_waves, _p(ticks), _p(due), _p(levels))
        if w < 0:
            raise RuntimeError("schedule rejected its inputs")
        return ticks[:w].copy(), due[:w].copy(), levels

    def schedule_adaptive(self, horizon, requested, edges, wave_requested, wave_passed):
        req = np.ascontiguousarray(requested, dtype=np.float64)
        e = np.ascontiguousarray(edges, dtype=np.uint32)
        wr = np.ascontiguousarray(wave_requested, dtype=np.float64)
        wp = np.ascontiguousarray(wave_passed, dtype=np.uint8)
        n, max_waves = len(req), wr.shape[0]
        ticks = np.empty(max_waves, np.uint64)
        due = np.empty(max_waves, np.int32)
        levels = np.empty((max_waves, n), np.int32)
        self.lib.orc_schedule_adaptive.restype = C.c_int
        self.lib.orc_schedule_adaptive.argtypes = [C.c_int, C.c_double] + [P]*2 + [C.c_int] + [P]*5
        w = self.lib.orc_schedule_adaptive(n, horizon, _p(req), _p(e), max_waves, _p(wr), _p(wp), _p(ticks), _p(due),
                                           _p(levels))
        if w < 0:
            raise RuntimeError("schedule rejected its inputs")
        return ticks[:w].copy(), due[:w].copy(), levels[:w].copy()

    # ---- operator ----
    def new_ctl(self):
        buf = np.zeros(self.lib.orc_ctl_size(), np.uint8)
        self.lib.orc_ctl_init(_p(buf))
        return buf

    def bath_rates(self, y, z, size, numr, active, nbr, measure, lambda_0, threads=0):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (y, z, size, numr)]
        act = np.ascontiguousarray(active, dtype=np.uint8)
        nb = np.ascontiguousarray(nbr, dtype=np.int32)
        me = np.ascontiguousarray(measure, dtype=np.float64)
        out = np.empty((len(a[0]), 4), np.float64)
        self.lib.orc_bath_rates(len(a[0]), *[_p(v) for v in a], _p(act), _p(nb), _p(me), lambda_0, _p(out), threads)
        return out

    def pair_rates(self, y, z, size, numr, active, nbr, measure, lambda_0, threads=0):
        """n x N_K pair terms at the bath start, slot order (0 for empty slots)."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (y, z, size, numr)]
        act = np.ascontiguousarray(active, dtype=np.uint8)
        nb = np.ascontiguousarray(nbr, dtype=np.int32)
        me = np.ascontiguousarray(measure, dtype=np.float64)
        out = np.empty((len(a[0]), self.info["n_k"]), np.float64)
        self.lib.orc_pair_rates(len(a[0]), *[_p(v) for v in a], _p(act), _p(nb), _p(me), lambda_0, _p(out), threads)
        return out

    def operator(self, y, z, size, numr, active, spatial, nbr, measure, rng, seed, duration, total_mass, ctl,
                 threads=0, diag=None):
        """Advance size/numr/rng in place; returns the stats dict. Raises on an operator error."""
        ys = np.ascontiguousarray(y, dtype=np.float64)
        zs = np.ascontiguousarray(z, dtype=np.float64)
        act = np.ascontiguousarray(active, dtype=np.uint8)
        sp = np.ascontiguousarray(spatial, dtype=np.int32)
        nb = np.ascontiguousarray(nbr, dtype=np.int32)
        me = np.ascontiguousarray(measure, dtype=np.float64)
        for arr, dt in ((size, np.float64), (numr, np.float64), (rng, np.uint64)):
            if arr.dtype != dt or not arr.flags.c_contiguous:
                raise ValueError("size, numr and rng are updated in place and must be contiguous")
        st = Stats()
        err = C.create_string_buffer(256)
        rc = self.lib.orc_operator(len(ys), _p(ys), _p(zs), _p(size), _p(numr), _p(act), _p(sp), _p(nb), _p(me),
                                   _p(rng), seed, duration, total_mass, _p(ctl), threads, C.byref(st), err, 256,
                                   diag.ptr if diag is not None else None)
        if rc:
            raise RuntimeError(err.value.decode())
        return {f: getattr(st, f) for f, _ in st._fields_}


class Diagnostics:
    """COL_DIAGNOSTICS bookkeeping of Oracle.operator: the controller archive (schema 2) since the last
    reset and the last operator's record (the schema-1 log line without wall-clock fields)."""

    def __init__(self, oracle: Oracle):
        self.lib = oracle.lib
        self.ptr = self.lib.orc_diag_new()

    def __del__(self):
        if getattr(self, "ptr", None):
            self.lib.orc_diag_free(self.ptr)
            self.ptr = None

    def reset(self) -> None:
        self.lib.orc_diag_reset(self.ptr)

    def _json(self, which: int) -> str:
        n = self.lib.orc_diag_json(self.ptr, which, None, 0)
        buf = C.create_string_buffer(n + 1)
        self.lib.orc_diag_json(self.ptr, which, buf, n + 1)
        return buf.value.decode()

    def controller(self) -> dict:
        import json
        return json.loads(self._json(0))

    def last_operator(self) -> dict:
        import json
        return json.loads(self._json(1))
