"""The oracle's k-d tree (left-balanced build by per-level tag sorts, closest-corner top-K query with
a bounded max-heap) pinned without the reference: every node separates its subtrees along its split
dimension, split dimensions are the widest extent of each subtree's domain (world box clipped by the
ancestors' planes, a later dimension winning ties), subtree sizes follow the left-balanced layout,
and each query's heap holds exactly the brute-force top-K under the (d^2, code) order."""
import numpy as np
import pytest


def _subtree(n, root):
    out, stack = [], [root]
    while stack:
        m = stack.pop()
        if m < n:
            out.append(m)
            stack += [2*m + 1, 2*m + 2]
    return out


def _domain(node, world, pts, idx, dim):
    lo, hi = world[:3].copy(), world[3:].copy()
    cur = node
    while cur > 0:
        par = (cur + 1)//2 - 1
        d = dim[par]
        split = pts[idx[par], d]
        if cur & 1:
            hi[d] = min(split, hi[d])
        else:
            lo[d] = max(split, lo[d])
        cur = par
    return lo, hi


def _widest(lo, hi):
    best, wb = 0, np.float32(0)
    for d in range(3):
        w = np.float32(hi[d] - lo[d])
        if w < wb:
            continue
        best, wb = d, w
    return best


def _cases():
    rng = np.random.default_rng(3)
    yield "uniform3d", rng.random((777, 3)).astype(np.float32)
    yield "flat2d", np.c_[rng.random((500, 2)), np.zeros(500)].astype(np.float32)
    g = np.array(np.meshgrid(np.arange(6), np.arange(5), np.arange(4))).reshape(3, -1).T
    yield "lattice_ties", (g*0.25).astype(np.float32)
    p = rng.random((300, 3)).astype(np.float32)
    yield "duplicates", np.r_[p, p[:100], p[:10]]
    yield "tiny", rng.random((5, 3)).astype(np.float32)
    yield "one", rng.random((1, 3)).astype(np.float32)


@pytest.mark.parametrize("name,pts", list(_cases()))
def test_kd_build_layout(oracles, name, pts):
    o = oracles["fiducial_3d"]
    n = len(pts)
    idx, dim, world = o.kd_build(pts)
    assert sorted(idx.tolist()) == list(range(n))
    assert np.array_equal(world[:3], pts.min(axis=0)) and np.array_equal(world[3:], pts.max(axis=0))
    assert dim[0] == _widest(world[:3], world[3:])
    for m in range(n):
        d = dim[m]
        split = pts[idx[m], d]
        left, right = _subtree(n, 2*m + 1), _subtree(n, 2*m + 2)
        assert all(pts[idx[j], d] <= split for j in left), (name, m)
        assert all(pts[idx[j], d] >= split for j in right), (name, m)
        if m > 0:
            lo, hi = _domain(m, world, pts, idx, dim)
            assert d == _widest(lo, hi), (name, m)
    # left-balanced: all levels full but the last, which fills from the left
    levels = int(np.floor(np.log2(n))) + 1 if n else 0
    assert n >= 2**(levels - 1)


@pytest.mark.parametrize("name,pts", [c for c in _cases() if c[0] != "one"])
@pytest.mark.parametrize("k", [1, 4, 16, 32])
def test_kd_query_matches_brute_force(oracles, name, pts, k):
    o = oracles["fiducial_3d"]
    n = len(pts)
    rng = np.random.default_rng(k)
    active = (rng.random(n) > 0.15).astype(np.uint8)
    q = np.full(n, 0.35, np.float32)
    q[::7] = 0.05                                           # some queries with fewer than K in range
    tree = o.kd_build(pts)
    qidx = np.arange(n, dtype=np.int32)
    codes, d2 = o.kd_query(pts, active, tree, qidx, k, q)
    bc, bd = o.topk_brute(pts, active, pts, q, k)
    for i in range(n):
        got = sorted((float(d2[i, s]), int(codes[i, s])) for s in range(k) if codes[i, s] >= 0)
        want = [(float(bd[i, s]), int(bc[i, s])) for s in range(k) if bc[i, s] >= 0]
        assert got == want, (name, k, i)
        # heap order: every parent key >= its children's
        key = [(np.float32(d2[i, s]).view(np.uint32).item() << 32) | (int(codes[i, s]) & 0xffffffff)
               for s in range(k)]
        for s in range(1, k):
            assert key[(s - 1)//2] >= key[s]
