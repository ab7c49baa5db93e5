"""Ensemble parity against the closed-form Smoluchowski solution (model smol_const), on the GPU.

The reference's constant-kernel campaign (val/paper/smoluchowski/README.md:123-138): unit monomers, the
constant synthetic kernel and a unit-volume pair measure make every run a sample of the normalised
Smoluchowski equation, whose solution is g(t) = 1 / (1 + t/2), number density M0 = g, second moment
M2 = 1 + t and mass fraction at mass k equal to k n_k = k g^2 (1 - g)^(k - 1).

Tolerances (the reference's scorers set none, guide_tests.md:769-773): the ensemble mean of M0 and of M2
lies within 4 standard errors of the seed-to-seed scatter plus 0.5% of the exact value (the frozen-bath
approximation at COL_BATH_EPS = 0.02; the CPU restatement of the same algorithm stays below 0.45% at 16k
representatives per realization); the pooled mass-fraction histogram passes a chi-squared test whose
per-bin variance adds a 2% systematic term to the multinomial one, at p > 1e-4. Both the plain operator
and the learned one (a majorant trained on this state, every realization thinning) are held to the
closed form, and to each other.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, SEEDS, TIMES = 65536, 8, (1.0, 2.0)


def _exact(t):
    g = 1.0/(1.0 + 0.5*t)
    return g, 1.0 + t


def _ensemble(gpu, majorant=None, safety=1.2):
    """SEEDS independent realizations advanced together (one batched call per output time); returns the
    masses per realization at every time and the summed statistics."""
    from gamedev_amd import CollisionOperator, evolve_batch, synth
    ops = []
    for seed in range(SEEDS):
        st = synth.smoluchowski_state(N, seed=seed)
        d = synth.to_device(st, gpu)
        op = CollisionOperator("smol_const", N, st["total_mass"], seed=seed + 1)
        op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
        op.geometry()
        if majorant is not None:
            op.use_majorant(majorant, safety, min_owners=1)
        ops.append((op, d, st["total_mass"]))
    out, t_prev, tot = {}, 0.0, {}
    for t in TIMES:
        print(f"[smol] {'learned' if majorant else 'plain'} ensemble to t={t}", flush=True)
        for s in evolve_batch([op for op, _, _ in ops], t - t_prev):
            for k, v in s.items():
                if isinstance(v, (int, float)):
                    tot[k] = tot.get(k, 0) + v
        t_prev = t
        rows = []
        for op, d, M in ops:
            s = d["size"].cpu().numpy()
            numr = d["numr"].cpu().numpy()
            m = s**3                                   # density 6/pi: grain mass = size^3 (monomer = 1)
            rows.append((numr.sum()/M, (numr*m*m).sum()/M, m))
        out[t] = rows
    return out, tot


def _check_moments(res, label):
    for t, rows in res.items():
        g, m2 = _exact(t)
        M0 = np.array([r[0] for r in rows])
        M2 = np.array([r[1] for r in rows])
        se0, se2 = M0.std(ddof=1)/math.sqrt(SEEDS), M2.std(ddof=1)/math.sqrt(SEEDS)
        print(f"[smol] {label} t={t}: M0 {M0.mean():.5f} (exact {g:.5f}, se {se0:.1e}), "
              f"M2 {M2.mean():.4f} (exact {m2:.4f}, se {se2:.1e})", flush=True)
        assert abs(M0.mean() - g) <= 4*se0 + 0.005*g, (t, M0.mean(), g)
        assert abs(M2.mean() - m2) <= 4*se2 + 0.005*m2, (t, M2.mean(), m2)


def _chi2(res, t, label):
    """Pooled mass-fraction histogram against k g^2 (1 - g)^(k-1): bins with at least 20 expected
    representatives, the rest merged into one tail bin."""
    from scipy.stats import chi2 as chi2_dist
    g, _ = _exact(t)
    m = np.concatenate([r[2] for r in res[t]])
    k = np.rint(m).astype(np.int64)
    assert np.all(np.abs(m - k) < 1e-9*m), "sticking of unit monomers keeps integer masses"
    total = len(k)
    kmax = 1
    while total*kmax*g*g*(1 - g)**(kmax - 1) >= 20:
        kmax += 1
    kk = np.arange(1, kmax)
    expected = total*kk*g*g*(1 - g)**(kk - 1)
    observed = np.array([(k == v).sum() for v in kk], float)
    expected = np.append(expected, total - expected.sum())
    observed = np.append(observed, (k >= kmax).sum())
    var = expected + (0.02*expected)**2
    stat = float(((observed - expected)**2/var).sum())
    dof = len(expected) - 1
    p = float(chi2_dist.sf(stat, dof))
    print(f"[smol] {label} t={t}: chi2 {stat:.1f} on {dof} dof, p = {p:.3g}; "
          f"fractions k=1..4 {observed[:4]/total} vs {expected[:4]/total}", flush=True)
    assert p > 1e-4
    return observed/total


def test_constant_kernel_plain_matches_closed_form(gpu):
    res, tot = _ensemble(gpu)
    assert tot["events"] > SEEDS*N*0.5
    _check_moments(res, "plain")
    _chi2(res, TIMES[-1], "plain")


def test_constant_kernel_learned_matches_closed_form_and_plain(gpu):
    """The learned operator on the same problem: bound trained on this state (the exact rate over S is
    one for the constant kernel), thinning in every realization."""
    import torch
    from gamedev_amd import CollisionOperator, Majorant, synth
    st = synth.smoluchowski_state(N, seed=100)
    d = synth.to_device(st, gpu)
    op = CollisionOperator("smol_const", N, st["total_mass"], seed=100)
    op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
    op.geometry()
    X, T = [], []
    for _ in range(3):
        f, S, R = op.majorant_sample()
        ok = (S > 0) & (R > 0)
        X.append(f[ok])
        T.append(torch.log(R[ok]/S[ok]).float())
        op.evolve(0.3)
    m = Majorant("smol_const", seed=2)
    losses, safety = m.fit(torch.cat(X), torch.cat(T), epochs=4, batch=8192, lr=1e-3, seed=2)
    print(f"[smol] majorant losses {losses[0]:.4g} -> {losses[-1]:.4g}, safety {safety:.3f}", flush=True)
    res, tot = _ensemble(gpu, m, safety)
    print(f"[smol] learned: candidates {tot['candidates']} accepted {tot['accepted']} checks "
          f"{tot['exact_checks']} violations {tot['violations']} thinned {tot['thinned']}", flush=True)
    assert tot["thinned"] > 0 and tot["candidates"] > 0
    _check_moments(res, "learned")
    f_l = _chi2(res, TIMES[-1], "learned")
    assert tot["violations"] <= 0.01*max(1, tot["exact_checks"])
    assert np.all(np.isfinite(f_l))
