"""Pins of the oracle's pair speed, Stokes number, vertical overlap and partner walk (CPU only).

Each check compares the oracle with an independent restatement (tests/physics_ref.py, written from the
reference's documented closed forms) or with hand-derived known answers. The planted-mistake tests
build the oracle with -DORC_PLANT=n (a dropped factor, a wrong cap, a sign, a wrong index, a wrong
comparison) and require the matching check to fail, so a plausible mistake of that kind cannot pass.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from physics_ref import Disk

HALF_PI = 0.5*math.pi
REL = 1e-11          # colphys probe tolerance of the reference is 2e-11 (validate_colphys.py:210-221)


def _close(got, ref, rel, what):
    if not (abs(got - ref) <= rel*abs(ref)):
        raise AssertionError(f"{what}: oracle {got!r} vs independent {ref!r} (rel {abs(got - ref)/abs(ref):.3e})")


def _probe_stokes(disk, R, Z):
    """The colphys Stokes probes (val/swarm/src/colphys_main.cuh:40-63) at the local Reynolds number: one
    point inside each Ormel-Cuzzi regime and both sides of every regime boundary."""
    c = disk.c
    column = c["SIGMA_0"]*(R/c["R_0"])**c["IDX_P"]*disk.stratification(R, Z)
    ris = (c["REYNOLDS_0"]*column/c["SIGMA_0"])**-0.5
    bounds = [0.2*ris, ris/1.6, 5*ris, 0.2, 1.0]
    return [0.02*ris, 0.4*ris, 2*ris, math.sqrt(5*ris*0.2), 0.5, 2.0] + \
        [b*f for b in bounds for f in (1 - 1e-6, 1 + 1e-6)]


def check_pair_speeds(o, model):
    """Unequal sizes (ratio 0.3 and pairs straddling the settling cap St = 0.5) at midplane and
    off-plane owners, at R on both sides of R_0."""
    disk = Disk(model)
    if disk.vertical:
        positions = [(1.0, HALF_PI), (1.7, HALF_PI - 0.06), (0.45, HALF_PI + 0.1)]
    else:
        positions = [(1.0, HALF_PI), (1.37, HALF_PI), (0.62, HALF_PI)]
    n = 0
    for y, z in positions:
        R, Z = disk.cylindrical(y, z)
        pairs = [(st, 0.3*st) for st in _probe_stokes(disk, R, Z)]
        pairs += [(0.8, 0.3), (0.45, 0.7), (3.0, 0.6), (0.55, 0.05), (0.2, 0.9)]
        for st_i, st_j in pairs:
            si, sj = disk.size_for_stokes(R, Z, st_i), disk.size_for_stokes(R, Z, st_j)
            _close(o.stokes(y, z, si), disk.stokes(R, Z, si), 1e-13, f"Stokes number at y={y}, z={z}")
            _close(o.vrel_pair(y, z, si, sj), disk.pair_speed(y, z, si, sj), REL,
                   f"pair speed at y={y}, z={z}, St {st_i:.3g}/{st_j:.3g}")
            _close(o.vrel_pair(y, z, sj, si), disk.pair_speed(y, z, sj, si), REL, "pair speed, swapped")
            n += 1
    assert n > 50


def check_overlap(o, model):
    """Vertically integrated physical rate with partners at other radii (R_i != R_j): the hand value of
    the overlap factor and the independent rate."""
    disk = Disk(model)
    c = disk.c
    assert c["N_Z"] == 1 and c["COAG_KERNEL"] == 3
    cases = [(0.8, 1.2, 0.3, 0.09), (1.4, 0.55, 0.7, 0.2), (1.0, 1.0, 2.0, 0.5), (0.6, 1.45, 0.05, 0.4)]
    for yi, yj, si, sj in cases:
        nj = 37.0
        rate, v = o.chain_rate(yi, HALF_PI, yj, HALF_PI, si, sj, nj, 3.0)
        # hand value: H_g = R h_g(R) with h_g = ASPR_0 (R / R_0)^((q + 1) / 2) at each particle's radius
        hi = yi*c["ASPR_0"]*(yi/c["R_0"])**((c["IDX_Q"] + 1)/2)
        hj = yj*c["ASPR_0"]*(yj/c["R_0"])**((c["IDX_Q"] + 1)/2)
        factor = rate/(nj*v*math.pi*(si + sj)**2/4.0)
        _close(factor, 1.0/math.sqrt(2.0*math.pi*(hi*hi + hj*hj)), 1e-13, f"overlap factor at R {yi}/{yj}")
        _close(rate, disk.physical_rate(yi, HALF_PI, yj, HALF_PI, si, sj, nj), REL, f"rate at R {yi}/{yj}")


def check_partner_walk(o):
    """Known answers of the chain's partner pick (col_chain_run.cu:299-320) and numpy searchsorted."""
    r = np.array([0.0, 1.0, 0.0, 2.0, 3.0])
    known = {0.0: 1, 0.999: 1, 1.0: 3, 2.5: 3, 3.0: 4, 5.999: 4, 6.0: 4, 6.5: 4}
    for target, slot in known.items():
        got = o.partner_walk(r, target)
        if got != slot:
            raise AssertionError(f"partner for target {target}: {got}, expected {slot}")
    # roundoff fallback: no running sum passes the target -> the last positive slot, not the first
    for rates, last in (([2.0, 3.0, 0.0, 0.0], 1), ([0.0, 4.0, 0.0, 1.0, 0.0], 3), ([5.0], 0)):
        got = o.partner_walk(np.array(rates), sum(rates))
        if got != last:
            raise AssertionError(f"fallback for {rates}: {got}, expected {last}")
    if o.partner_walk(np.zeros(4), 0.0) != -1:
        raise AssertionError("no positive slot must give -1 (chain error 5)")
    rng = np.random.default_rng(7)
    for _ in range(2000):
        k = int(rng.integers(1, 65))
        rates = rng.exponential(size=k)*(rng.uniform(size=k) < 0.7)
        if not rates.any():
            continue
        cs = np.cumsum(rates)
        target = rng.uniform()*cs[-1]
        want = int(np.searchsorted(cs, target, side="right"))
        if want == k:
            want = int(np.nonzero(rates > 0)[0][-1])
        got = o.partner_walk(rates, target)
        if got != want:
            raise AssertionError(f"partner {got} != searchsorted {want}")


# ---- the oracle as built ----
@pytest.mark.parametrize("model", ["fiducial_3d", "wedge_2d"])
def test_pair_speed_unequal_sizes(oracles, model):
    check_pair_speeds(oracles[model], model)


def test_vertical_overlap_unequal_radii(oracles):
    check_overlap(oracles["wedge_2d"], "wedge_2d")


def test_partner_walk_known_answers(oracles):
    check_partner_walk(oracles["fiducial_3d"])


def test_partner_walk_frequencies(oracles):
    """Slot k is chosen for a fraction r_k / R of evenly spaced targets."""
    o = oracles["fiducial_3d"]
    r = np.array([0.5, 0.0, 1.5, 0.25, 0.0, 1.75])
    m = 40000
    hits = np.zeros(len(r), int)
    for t in (np.arange(m) + 0.5)/m*r.sum():
        hits[o.partner_walk(r, t)] += 1
    np.testing.assert_allclose(hits/m, r/r.sum(), atol=2.0/m)


# ---- planted mistakes: each must be rejected by its pin ----
PLANTS = {
    1: ("fiducial_3d", "drift: factor 2 of dv_R dropped", lambda o: check_pair_speeds(o, "fiducial_3d")),
    2: ("fiducial_3d", "settling cap St = 1 instead of 0.5", lambda o: check_pair_speeds(o, "fiducial_3d")),
    3: ("wedge_2d", "overlap: partner scale height at the owner's radius", lambda o: check_overlap(o, "wedge_2d")),
    4: ("fiducial_3d", "Stokes number: sign of the radial exponent", lambda o: check_pair_speeds(o, "fiducial_3d")),
    5: ("fiducial_3d", "partner walk: >= instead of >", check_partner_walk),
    6: ("fiducial_3d", "partner walk: fallback to the first positive slot", check_partner_walk),
    7: ("fiducial_3d", "rotation support: sign of the q (1 - u) term", lambda o: check_pair_speeds(o, "fiducial_3d")),
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_is_rejected(tmp_path, plant):
    from oracle.build import build_planted
    from oracle.pyoracle import Oracle
    model, what, check = PLANTS[plant]
    o = Oracle(model, path=build_planted(model, plant, tmp_path))
    with pytest.raises(AssertionError):
        check(o)
