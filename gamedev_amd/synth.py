"""Synthetic swarms for the benchmark and parity workloads (host numpy, particle order).

plain_2d_state:     65,536 representatives, vertically integrated disk (N_Z = 1), r uniform on [0.5, 2.0)
                    (equal weights with surface density ~ 1/r give dN/dr = const), phi uniform,
                    monodisperse size 1e-4, equal weights, total mass 1e-4.
fiducial_3d_state:  1e7 representatives by default, full 3D, r uniform on [0.3, 3.0), polar angle
                    Gaussian about pi/2 with h/r = 0.05 r^0.25, log-normal sizes (median 1e-4,
                    sigma_ln = 1, clipped to [1e-5, 1e-1]), equal-mass representatives, total dust
                    mass = 0.01 x the gas mass 2 pi Sigma_0 (3.0 - 0.3).
smoluchowski_state: the constant-kernel Smoluchowski benchmark: unit monomers of unit mass on an
                    annulus, unit-volume measure (model smol_const).
clumped_3d_state:   the clumped stress case: the fiducial background with a vortex overdensity (x10,
                    width 0.05 r, at r = 1, phi = pi) and a pressure-bump ring (x5, width 0.03, at
                    r = 1.5), sizes log-normal with sigma_ln = 1.5.
Velocities (Keplerian plus dispersion) do not enter any collision rate: the reference's relative speed
is query-local and analytic (inc/swarm/_collision.cuh:275-325), so the state holds none.
The size variable is the reference's par_size and takes the drawn value.
"""
from __future__ import annotations

import math

import numpy as np

RHO_0 = 1.0


def grain_mass(size, rho=RHO_0):
    return math.pi*rho*size**3/6.0


def plain_2d_state(n: int = 65536, seed: int = 0, total_mass: float = 1e-4, size: float = 1e-4) -> dict:
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.uniform(0.5, 2.0, n)
    phi = rng.uniform(-math.pi, math.pi, n)
    z = np.full(n, 0.5*math.pi)
    s = np.full(n, size)
    w = total_mass/n
    numr = w/grain_mass(s)
    return {"x": phi, "y": r, "z": z, "size": s, "numr": numr, "total_mass": total_mass}


def wedge_2d_state(n: int = 65536, seed: int = 0, r_min: float = 0.5, r_max: float = 1.5, x_min: float = -0.1,
                   x_max: float = 0.1, s_min: float = 0.05, s_max: float = 6.4) -> dict:
    """Vertically integrated azimuthal wedge: r uniform, azimuth uniform over the wedge, log-uniform sizes,
    equal-mass representatives carrying 1% of a unit-surface-density gas wedge."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.uniform(r_min, r_max, n)
    phi = rng.uniform(x_min, x_max, n)
    z = np.full(n, 0.5*math.pi)
    s = np.exp(rng.uniform(math.log(s_min), math.log(s_max), n))
    total_mass = 0.01*0.5*(x_max - x_min)*(r_max**2 - r_min**2)
    w = total_mass/n
    numr = w/grain_mass(s)
    return {"x": phi, "y": r, "z": z, "size": s, "numr": numr, "total_mass": total_mass}


def fiducial_3d_state(n: int = 10_000_000, seed: int = 0, r_min: float = 0.3, r_max: float = 3.0,
                      sigma_ln: float = 1.0) -> dict:
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.uniform(r_min, r_max, n)
    phi = rng.uniform(-math.pi, math.pi, n)
    aspect = 0.05*r**0.25
    theta = 0.5*math.pi + aspect*rng.standard_normal(n)
    s = np.clip(1e-4*np.exp(sigma_ln*rng.standard_normal(n)), 1e-5, 1e-1)
    gas_mass = 2.0*math.pi*1e-2*(r_max - r_min)
    total_mass = 0.01*gas_mass
    w = total_mass/n
    numr = w/grain_mass(s)
    return {"x": phi, "y": r, "z": theta, "size": s, "numr": numr, "total_mass": total_mass}


def clumped_3d_state(n: int, seed: int = 0, r_min: float = 0.3, r_max: float = 3.0, sigma_ln: float = 1.5,
                     vortex=(1.0, math.pi, 0.05, 10.0), ring=(1.5, 0.03, 5.0)) -> dict:
    """The fiducial 3D background with a Gaussian azimuthal vortex overdensity (centre r, phi, width as a
    fraction of r, peak factor) and a radial pressure-bump ring (radius, width, peak factor), equal-mass
    representatives placed by rejection from the background, log-normal sizes of the given spread."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rv, pv, wv, av = vortex
    rr, wr, ar = ring
    peak = max(av, ar)
    r = np.empty(0)
    phi = np.empty(0)
    while len(r) < n:
        m = 2*(n - len(r)) + 1024
        rc = rng.uniform(r_min, r_max, m)
        pc = rng.uniform(-math.pi, math.pi, m)
        dphi = np.angle(np.exp(1j*(pc - pv)))
        w = wv*rv
        gv = np.exp(-((rc - rv)**2 + (rv*dphi)**2)/(2*w*w))
        gr = np.exp(-(rc - rr)**2/(2*wr*wr))
        f = 1.0 + (av - 1.0)*gv + (ar - 1.0)*gr
        keep = rng.uniform(0.0, peak, m) < f
        r = np.concatenate([r, rc[keep]])
        phi = np.concatenate([phi, pc[keep]])
    r, phi = r[:n], phi[:n]
    aspect = 0.05*r**0.25
    theta = 0.5*math.pi + aspect*rng.standard_normal(n)
    s = np.clip(1e-4*np.exp(sigma_ln*rng.standard_normal(n)), 1e-5, 1e-1)
    gas_mass = 2.0*math.pi*1e-2*(r_max - r_min)
    total_mass = 0.01*gas_mass
    w = total_mass/n
    numr = w/grain_mass(s)
    return {"x": phi, "y": r, "z": theta, "size": s, "numr": numr, "total_mass": total_mass}


def smoluchowski_state(n: int = 65536, seed: int = 0, total_mass: float = 1e30) -> dict:
    """Constant-kernel Smoluchowski benchmark (model smol_const): unit monomers (diameter 1, density 6/pi, so
    mass 1) on a full-azimuth annulus 0.5 <= r < 1.5, equal represented mass total_mass / n
    (val/paper/smoluchowski/README.md: BENCHMARK_MASS = 1e30; positions uniform instead of the campaign's
    jittered ring grid, which the unit-volume measure makes irrelevant to the rates)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.uniform(0.5, 1.5, n)
    phi = rng.uniform(-math.pi, math.pi, n)
    z = np.full(n, 0.5*math.pi)
    s = np.ones(n)
    numr = np.full(n, total_mass/n)
    return {"x": phi, "y": r, "z": z, "size": s, "numr": numr, "total_mass": total_mass}


def ensemble_sizes(count: int, lo: float, hi: float, seed: int = 12345, total: int | None = None) -> list[int]:
    """Realization sizes floor(2^U[lo, hi]); with `total`, rescaled so that they sum to it."""
    rng = np.random.default_rng(seed)
    sizes = [int(2.0**rng.uniform(lo, hi)) for _ in range(count)]
    if total:
        scale = total/sum(sizes)
        sizes = [max(64, int(round(v*scale))) for v in sizes]
        sizes[-1] += total - sum(sizes)
    return sizes


def to_device(state: dict, device="cuda:0") -> dict:
    import torch
    return {k: torch.as_tensor(np.ascontiguousarray(v), dtype=torch.float64, device=device)
            for k, v in state.items() if isinstance(v, np.ndarray)}
