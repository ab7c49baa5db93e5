"""GPU parity of the geometry epoch against the CPU oracle (bit-exact sites, keys, sort order and
neighbour lists; KNN measure to 1e-12 relative). Calls the HIP library through the C ABI."""
import math

import numpy as np
import pytest

from helpers import oracle_geometry, root_cube

pytestmark = pytest.mark.gpu


def _make(model, state, gpu, seed=0):
    from gamedev_amd import CollisionOperator, synth
    dev = synth.to_device(state, gpu)
    op = CollisionOperator(model, len(state["x"]), state["total_mass"], seed=seed)
    op.bind_soa(dev["x"], dev["y"], dev["z"], dev["size"], dev["numr"])
    op.geometry()
    return op, dev


def _states():
    from gamedev_amd import synth
    return {
        "plain_2d": synth.plain_2d_state(65536, seed=0),
        "fiducial_3d": synth.fiducial_3d_state(60000, seed=1),
    }


@pytest.fixture(scope="module")
def built(gpu):
    out = {}
    for model, st in _states().items():
        print(f"[geometry] building {model} n={len(st['x'])}", flush=True)
        op, dev = _make(model, st, gpu)
        out[model] = (op, dev, st)
    return out


@pytest.mark.parametrize("model", ["plain_2d", "fiducial_3d"])
def test_sites_match_oracle(built, oracles, model):
    op, _, st = built[model]
    o = oracles[model]
    pts, act, grp = op.sites()
    opts, _, oq, oact, bad = o.site_init(st["x"], st["y"], st["z"], st["size"], st["numr"])
    assert bad == 0
    assert np.array_equal(act, oact)
    assert np.array_equal(grp.astype(np.int32), o.space_bin(st["x"], st["y"], st["z"]))
    # FP64 cos/sin/pow come from different libraries: the float results agree except in rare last-ulp cases
    exact = np.all(pts[:, :3] == opts, axis=1) & (pts[:, 3] == oq)
    assert exact.mean() > 0.999
    ulps = np.abs(pts[:, :3].view(np.int32) - opts.view(np.int32))
    assert ulps.max() <= 1


@pytest.mark.parametrize("model", ["plain_2d", "fiducial_3d"])
def test_keys_and_sort_bit_exact(built, oracles, model):
    op, _, st = built[model]
    o = oracles[model]
    pts, act, _ = op.sites()
    keys, idx = op.morton()
    org, w = root_cube(o.info)
    okeys = o.morton_keys(pts[:, :3], org, w, o.info["morton_max_level"], o.info["dim"])
    operm = o.sort_perm(okeys)
    assert np.array_equal(idx, operm)
    assert np.array_equal(keys, okeys[operm])


@pytest.mark.parametrize("model", ["plain_2d", "fiducial_3d"])
def test_neighbors_bit_exact(built, oracles, model):
    op, _, st = built[model]
    o = oracles[model]
    pts, act, _ = op.sites()
    codes, meas = op.neighbors()
    ref = oracle_geometry(o, pts[:, :3], act, pts[:, 3], st["y"], st["z"])
    assert np.array_equal(codes, ref["codes"])
    # independent brute force on a query subset
    sub = np.arange(0, len(pts), 97)
    sub = sub[act[sub] == 1]
    bc, _ = o.topk_brute(pts[:, :3], act, pts[sub, :3], pts[sub, 3], o.info["n_k"])
    assert np.array_equal(codes[sub], bc)
    ok = act == 1
    a, b = meas[ok], ref["measure"][ok]
    # an owner whose only neighbour is itself has radius 0 and measure 0 on both sides
    assert np.array_equal(a == 0, b == 0)
    nz = b != 0
    rel = np.abs(a[nz] - b[nz]) / np.abs(b[nz])
    assert rel.max() <= 1e-12
    assert np.all(meas[~ok] == 0)


@pytest.mark.parametrize("model", ["plain_2d", "fiducial_3d"])
def test_dependency_graph(built, oracles, model):
    op, _, st = built[model]
    pts, act, grp = op.sites()
    codes, _ = op.neighbors()
    rows = op.graph()
    expect = np.zeros_like(rows)
    for i in np.nonzero(act)[0]:
        for c in codes[i]:
            if c >= 0:
                expect[grp[i]] |= np.uint64(1) << np.uint64(grp[c // 3])
    assert np.array_equal(rows, expect)


@pytest.mark.parametrize("model", ["plain_2d", "fiducial_3d"])
def test_bath_rates_match_oracle(built, oracles, model):
    op, dev, st = built[model]
    o = oracles[model]
    op.rate_pass()
    cached = op.cached()
    pts, act, _ = op.sites()
    codes, meas = op.neighbors()
    lam = len(st["x"]) / o.info["n_k"] / st["total_mass"]
    ref = o.bath_rates(st["y"], st["z"], st["size"], st["numr"], act, codes, meas, lam)
    ok = act == 1
    for f in range(4):
        a, b = cached[ok, f], ref[ok, f]
        scale = np.maximum(np.abs(b), 1e-300)
        assert np.max(np.abs(a - b) / scale) <= 1e-12, f"field {f}"


@pytest.mark.parametrize("model", ["plain_2d", "fiducial_3d"])
def test_pair_rates_match_oracle(built, oracles, model):
    """Every owner's K pair rates (gdx_col_pair_rates) against the oracle's pair terms, 1e-12 relative per
    pair; empty slots and inactive owners are exactly zero on both sides."""
    op, dev, st = built[model]
    o = oracles[model]
    rates = op.pair_rates()
    pts, act, _ = op.sites()
    codes, meas = op.neighbors()
    lam = len(st["x"]) / o.info["n_k"] / st["total_mass"]
    ref = o.pair_rates(st["y"], st["z"], st["size"], st["numr"], act, codes, meas, lam)
    assert np.array_equal(rates == 0.0, ref == 0.0)
    nz = ref != 0.0
    assert nz.sum() > 0.5*act.sum()*o.info["n_k"]
    rel = np.abs(rates[nz] - ref[nz])/np.abs(ref[nz])
    print(f"[pair rates] {model}: {nz.sum()} nonzero pairs, max rel {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-12
    # the owner totals of the rate pass are the sums of these terms
    tot = op.cached()[act == 1, 0]
    s = rates[act == 1].sum(axis=1)
    assert np.allclose(tot, s, rtol=1e-13, atol=0.0)
