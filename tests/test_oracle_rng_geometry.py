"""CPU pins of the oracle's random numbers and geometry (no GPU)."""
import math

import numpy as np
import pytest

from helpers import root_cube


# ---- Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11) ----
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff]*4, [0xffffffff]*2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(oracles, ctr, key, expect):
    assert oracles["plain_2d"].philox(ctr, key) == expect


def test_uniform_range_and_distribution(oracles):
    from scipy import stats
    o = oracles["plain_2d"]
    u = np.array([o.uniform(i % 97, 12345, i // 97) for i in range(20000)])
    assert u.min() > 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3


# ---- Morton keys: bit-by-bit interleave of the production quantisation ----
def _key_loop(p, origin, width, level, dim):
    cells = 1 << level
    scale = np.float32(cells)/np.float32(width)
    ijk = []
    for a in range(3):
        if a == 2 and dim == 2:
            ijk.append(0)
            continue
        v = int((np.float32(p[a]) - np.float32(origin[a]))*scale)
        ijk.append(min(cells - 1, max(0, v)))
    key = 0
    for bit in range(level):
        for a in range(3):
            key |= ((ijk[a] >> bit) & 1) << (3*bit + a)
    return key


@pytest.mark.parametrize("dim", [2, 3])
def test_morton_keys_bit_loop(oracles, dim):
    o = oracles["fiducial_3d"]
    rng = np.random.default_rng(5)
    pts = rng.uniform(-3.0, 3.0, (500, 3)).astype(np.float32)
    pts[:10] = [[-3.0003, -3.0003, -3.0003]]*10   # clamped at the lower root corner
    org, w = root_cube(o.info)
    keys = o.morton_keys(pts, org, w, 20, dim)
    for i in range(len(pts)):
        assert int(keys[i]) == _key_loop(pts[i], org, w, 20, dim)


def test_sort_is_stable_and_sorted(oracles):
    o = oracles["plain_2d"]
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 50, 5000).astype(np.uint64)
    perm = o.sort_perm(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))


# ---- hierarchy: independent recursive linear partition (val/swarm/mod/test_knn/topology.cuh:70-110) ----
def _recursive(keys_sorted, lo, hi, dim, target, depth, lower, width, out, level=0):
    nid = len(out)
    out.append({"begin": lo, "count": hi - lo, "lower": tuple(lower), "width": width, "children": {}})
    if hi - lo <= target or level == depth:
        return nid
    shift = 3*(depth - level - 1)
    cur = lo
    for code in range(1 << dim):
        first = cur
        while cur < hi and (int(keys_sorted[cur]) >> shift) & 7 == code:
            cur += 1
        if cur == first:
            continue
        half = np.float32(0.5)*np.float32(width)
        cl = [np.float32(lower[0]), np.float32(lower[1]), np.float32(lower[2])]
        if code & 1: cl[0] = np.float32(cl[0] + half)
        if code & 2: cl[1] = np.float32(cl[1] + half)
        if dim == 3 and code & 4: cl[2] = np.float32(cl[2] + half)
        out[nid]["children"][code] = _recursive(keys_sorted, first, cur, dim, target, depth, cl, half, out, level + 1)
    assert cur == hi
    return nid


def _topology_cases():
    rng = np.random.default_rng(761)
    cases = []
    for dim in (2, 3):
        base = np.stack([rng.integers(0, 4096, 513)/4096, rng.integers(0, 4096, 513)/4096,
                         rng.integers(0, 4096, 513)/4096 if dim == 3 else np.zeros(513)], 1)
        pts = np.concatenate([base, np.tile([0.5, 0.5, 0.5 if dim == 3 else 0.0], (300, 1)),
                              [[0, 0, 0]], [[1, 1, 1 if dim == 3 else 0]]]).astype(np.float32)
        for target in (1, 128):
            for depth in (1, 5, 20):
                cases.append((dim, target, depth, pts))
                cases.append((dim, target, depth, np.tile(np.float32([0.25, 0.25, 0.25]), (400, 1))))
                cases.append((dim, target, depth, np.zeros((1, 3), np.float32)))
    return cases


@pytest.mark.parametrize("case", range(36))
def test_hierarchy_matches_recursive_split(oracles, case):
    dim, target, depth, pts = _topology_cases()[case]
    o = oracles["fiducial_3d"]
    keys = o.morton_keys(pts, (0.0, 0.0, 0.0), 1.0, depth, dim)
    perm = o.sort_perm(keys)
    H = o.hierarchy(keys[perm], (0.0, 0.0, 0.0), 1.0, dim, target, depth)
    expect = []
    _recursive(keys[perm], 0, len(pts), dim, target, depth, [0.0, 0.0, 0.0], np.float32(1.0), expect)
    assert len(H["width"]) == len(expect)
    seen = set()

    def cmp(a, b):
        assert a not in seen
        seen.add(a)
        e = expect[b]
        assert H["begin"][a] == e["begin"] and H["count"][a] == e["count"]
        assert tuple(np.float32(v) for v in H["lower"][a]) == tuple(np.float32(v) for v in e["lower"])
        assert np.float32(H["width"][a]) == np.float32(e["width"])
        assert H["child_count"][a] == len(e["children"])
        for code in range(8):
            if code in e["children"]:
                cmp(H["child"][a][code], e["children"][code])
            else:
                assert H["child"][a][code] == -1
    cmp(0, 0)
    assert len(seen) == len(expect)


# ---- exact top-K: hand-derived answers of the reference's edge cases (knn_edge_tests.cu:95-210) ----
ANCHORS = {2: [[-2, -2, 0], [2, 2, 0]], 3: [[-2, -2, -2], [2, 2, 2]]}


def _brute(o, pts, q, k, active=None, queries=((0.0, 0.0, 0.0),)):
    pts = np.asarray(pts, np.float32)
    act = np.ones(len(pts), np.uint8) if active is None else np.asarray(active, np.uint8)
    qp = np.asarray(queries, np.float32)
    return o.topk_brute(pts, act, qp, np.full(len(qp), q, np.float32), k)


@pytest.mark.parametrize("dim", [2, 3])
def test_knn_equal_distance_ties(oracles, dim):
    pts = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]] + ([[0, 0, 1], [0, 0, -1]] if dim == 3 else []) + ANCHORS[dim]
    c, d = _brute(oracles["plain_2d"], pts, 1.0, 2)
    assert c[0].tolist() == [0, 3] and d[0].tolist() == [1.0, 1.0]


@pytest.mark.parametrize("dim", [2, 3])
def test_knn_coincident(oracles, dim):
    pts = [[0, 0, 0]]*300 + ANCHORS[dim]
    c, d = _brute(oracles["plain_2d"], pts, 0.25, 128)
    assert c[0].tolist() == [3*i for i in range(128)] and np.all(d == 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_knn_radius_boundary(oracles, dim):
    pts = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [np.nextafter(np.float32(1), np.float32(2)), 0, 0]] + ANCHORS[dim]
    c, _ = _brute(oracles["plain_2d"], pts, 1.0, 4)
    assert c[0].tolist() == [0, 3, 6, -1]


@pytest.mark.parametrize("dim", [2, 3])
def test_knn_fewer_than_k(oracles, dim):
    pts = [[0.25, 0, 0], [0, 0.5, 0]] + ANCHORS[dim]
    c, d = _brute(oracles["plain_2d"], pts, 0.75, 4)
    assert c[0].tolist() == [0, 3, -1, -1] and d[0][:2].tolist() == [0.0625, 0.25]


@pytest.mark.parametrize("dim", [2, 3])
def test_knn_split_planes(oracles, dim):
    vals = [-1.0, -0.5, 0.0, 0.5, 1.0]
    pts = [[x, y, z] for x in vals for y in vals for z in (vals if dim == 3 else [0.0])]
    c, d = _brute(oracles["plain_2d"], pts, 0.9, 4)
    centre = pts.index([0.0, 0.0, 0.0])
    ring = sorted(3*i for i, p in enumerate(pts) if abs(sum(v*v for v in p) - 0.25) < 1e-9)
    assert c[0].tolist() == [3*centre] + ring[:3]


def test_knn_radial_line(oracles):
    pts = [[0.5 + 0.25*i, 0, 0] for i in range(5)]
    c, _ = _brute(oracles["plain_2d"], pts, 0.6, 2, queries=[[0.5, 0, 0], [1.0, 0, 0], [1.5, 0, 0]])
    assert c.tolist() == [[0, 3], [6, 3], [12, 9]]


def test_knn_inactive_filter(oracles):
    pts = [[0.1*i, 0, 0] for i in range(1, 8)]
    active = [1 if i > 4 else 0 for i in range(1, 8)]
    c, _ = _brute(oracles["plain_2d"], pts, 1.0, 2, active=active)
    assert c[0].tolist() == [12, 15]


@pytest.mark.parametrize("dim", [2, 3])
def test_topk_brute_matches_numpy(oracles, dim):
    o = oracles["plain_2d"]
    rng = np.random.default_rng(dim)
    pts = np.round(rng.uniform(-1, 1, (600, 3))*32)/32          # many exact distance ties
    if dim == 2:
        pts[:, 2] = 0
    pts = pts.astype(np.float32)
    act = (rng.uniform(size=600) > 0.1).astype(np.uint8)
    q = np.float32(0.3)
    codes, d2 = o.topk_brute(pts, act, pts[:50], np.full(50, q, np.float32), 16)
    for i in range(50):
        dx, dy, dz = (pts[i] - pts).T
        d = (dx*dx + dy*dy) + dz*dz
        ok = (act == 1) & (d <= q*q)
        cand = sorted(zip(d[ok].tolist(), (3*np.nonzero(ok)[0]).tolist()))[:16]
        exp = [c for _, c in cand] + [-1]*(16 - len(cand))
        assert codes[i].tolist() == exp


@pytest.mark.parametrize("dim", [2, 3])
def test_topk_tree_equals_brute(oracles, dim):
    o = oracles["plain_2d"]
    rng = np.random.default_rng(10 + dim)
    n = 3000
    pts = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    pts[:200] = pts[200:400]                                    # coincident records
    if dim == 2:
        pts[:, 2] = 0
    act = (rng.uniform(size=n) > 0.05).astype(np.uint8)
    q = rng.uniform(0.02, 0.3, n).astype(np.float32)
    org, w = (-4.0, -4.0, -4.0), 8.0
    keys = o.morton_keys(pts, org, w, 20, dim)
    perm = o.sort_perm(keys)
    H = o.hierarchy(keys[perm], org, w, dim, 16, 20)
    codes, d2 = o.topk_tree(pts, act, perm, H, dim, 20, np.arange(n, dtype=np.int32), 16, q)
    bc, bd = o.topk_brute(pts, act, pts, q, 16)
    assert np.array_equal(codes, bc) and np.array_equal(d2, bd)
