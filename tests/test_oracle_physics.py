"""CPU pins of the oracle's pair physics, outcomes and KNN measure (no GPU)."""
import math

import numpy as np
import pytest

from physics_ref import oc07_speed_sq


# ---- KNN ball measure: closed forms and Monte-Carlo cap fractions ----
def test_ball_measure_interior_closed_forms(oracles):
    f3 = oracles["fiducial_3d"]
    a = 0.01
    assert f3.ball_measure(1.5, 0.5*math.pi, a) == pytest.approx(4.0*math.pi*a**3/3.0, rel=1e-15)
    p2 = oracles["plain_2d"]
    assert p2.ball_measure(1.2, 0.5*math.pi, a) == pytest.approx(math.pi*a*a, rel=1e-15)
    assert f3.ball_measure(1.0, 0.5*math.pi, 0.0) == 0.0


@pytest.mark.parametrize("model,dim", [("fiducial_3d", 3), ("plain_2d", 2)])
def test_ball_measure_single_face_monte_carlo(oracles, model, dim):
    """A ball crossing only the inner radial face keeps the fraction of its volume inside the half-space."""
    o = oracles[model]
    a, d = 0.05, 0.02
    y = o.info["y_min"] + d
    m = o.ball_measure(y, 0.5*math.pi, a)
    full = 4.0*math.pi*a**3/3.0 if dim == 3 else math.pi*a*a
    rng = np.random.default_rng(0)
    v = rng.normal(size=(400000, dim))
    v *= (rng.uniform(size=(400000, 1))**(1.0/dim))/np.linalg.norm(v, axis=1, keepdims=True)
    frac = np.mean(v[:, 0]*a > -d)
    assert m/full == pytest.approx(frac, abs=4*math.sqrt(frac*(1 - frac)/400000))


# ---- Ormel & Cuzzi (2007) regimes, evaluated independently in stopping-time form ----
def test_vrel_t_regimes_and_boundaries(oracles):
    """The colphys probe points (val/swarm/src/colphys_main.cuh:40-63): six regimes and ten straddles."""
    o = oracles["fiducial_3d"]
    R = 1.0
    h = 0.05*R**0.25
    sig = 1e-2
    alpha = 1e-3
    cs = h*R**-0.5
    vg2 = 1.5*alpha*cs*cs
    ris = 1.0/math.sqrt(1e8*(alpha/1e-3)*(sig/1e-2))
    bounds = [0.2*ris, ris/1.6, 5*ris, 0.2, 1.0]
    probes = [0.02*ris, 0.4*ris, 2*ris, math.sqrt(5*ris*0.2), 0.5, 2.0] + \
             [b*f for b in bounds for f in (1 - 1e-6, 1 + 1e-6)]
    for sl in probes:
        got = o.vrel_t(R, sl, 0.3*sl, h, sig)
        ref = math.sqrt(oc07_speed_sq(sl, 0.3*sl, 1.0, ris, vg2))
        assert got == pytest.approx(ref, rel=1e-11), sl


def test_vrel_t_equal_stokes_vanishes_when_coupled(oracles):
    o = oracles["fiducial_3d"]
    ris = 1e-4
    for st in (0.05*ris, 0.5*ris):
        assert o.vrel_t(1.0, st, st, 0.05, 1e-2) == 0.0


def test_vrel_pair_equal_sizes_zero(oracles):
    """Equal sizes: drift, settling and coupled turbulence all cancel (code units, no Brownian term)."""
    o = oracles["fiducial_3d"]
    assert o.vrel_pair(1.0, 0.5*math.pi + 0.03, 1e-6, 1e-6) == 0.0


def test_physical_rate_composition(oracles):
    o = oracles["fiducial_3d"]
    y, z, si, sj, nj = 1.3, 0.5*math.pi - 0.02, 3e-4, 1e-4, 250.0
    r, v = o.chain_rate(y, z, 1.31, z, si, sj, nj, 7.0)
    assert v == o.vrel_pair(y, z, si, sj)
    assert r == pytest.approx(nj*v*math.pi*(si + sj)**2/4.0, rel=1e-15)


def test_constant_kernel_rate(oracles):
    r, v = oracles["plain_2d"].chain_rate(1.0, 0.0, 1.1, 0.0, 1e-4, 1e-4, 123.0, 4.5)
    assert r == 4.5*123.0 and v == 0.0


# ---- outcome jump moments: expectations of the outcome sampler over its uniform ----
def _quad_moments(o, si, sj, high, npts=200001):
    u = (np.arange(npts) + 0.5)/npts
    j = np.empty(npts)
    for i, ui in enumerate(u):
        s, _, _ = o.sample_outcome(si, sj, high, float(ui))
        j[i] = abs(math.log(s/si))
    return j.mean(), (j*j).mean(), j.max()


@pytest.mark.parametrize("si,sj", [(1e-3, 5e-4), (1e-3, 3e-4), (5e-2, 1e-4)])
def test_erosion_moments_match_sampler(oracles, si, sj):
    o = oracles["fiducial_3d"]
    f, mean, second, mx = o.outcome_moments(si, sj, True)
    qm, qs, qx = _quad_moments(o, si, sj, True, 40001)
    assert mean == pytest.approx(qm, rel=1e-4)
    assert second == pytest.approx(qs, rel=1e-4)
    assert mx == pytest.approx(qx, rel=1e-3)


def test_catastrophic_moments_match_sampler(oracles):
    o = oracles["fiducial_3d"]
    si = 2e-3
    f, mean, second, mx = o.outcome_moments(si, si, True)
    assert f == 1.0
    qm, qs, _ = _quad_moments(o, si, si, True, 200001)
    assert mean == pytest.approx(qm, rel=1e-5)
    assert second == pytest.approx(qs, rel=1e-5)
    assert mx == pytest.approx(math.log(si/o.info["init_smin"]), rel=1e-12)


def test_catastrophic_series_branch_continuous(oracles):
    """The small-L series and the closed form agree where they meet (L = 1e-3)."""
    o = oracles["fiducial_3d"]
    smin = o.info["init_smin"]
    lo = o.outcome_moments(smin*math.exp(2*(1e-3 - 1e-13)), 1.0, True)
    hi = o.outcome_moments(smin*math.exp(2*(1e-3 + 1e-13)), 1.0, True)
    assert lo[1] == pytest.approx(hi[1], rel=1e-9) and lo[2] == pytest.approx(hi[2], rel=1e-7)


@pytest.mark.parametrize("q", [1.0, 1e-3, 1e-7])
def test_sticking_moments_and_packets(oracles, q):
    o = oracles["fiducial_3d"]
    si = 1e-3
    sj = si*q**(1/3)
    f, mean, second, mx = o.outcome_moments(si, sj, False)
    s, cat, lm = o.sample_outcome(si, sj, False, 0.5)
    r = sj/si
    qq = r*r*r
    G = max(1.0, math.floor(1e-4/qq)) if qq <= 1e-6 else 1.0
    assert f == pytest.approx(1.0/G, rel=1e-15)
    # log(s/si) of a ratio near one loses about -log10(G q) digits
    assert mean == pytest.approx(abs(math.log(s/si)), rel=1e-9)
    assert second == pytest.approx(mean*mean, rel=1e-15)
    assert s**3 == pytest.approx(si**3 + G*sj**3, rel=1e-14)
