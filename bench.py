#!/usr/bin/env python3
"""Collision-operator benchmark on one MI355X.

Workload (the fiducial model directory gamedev_amd/mod/fiducial_3d): 1e7 representatives in a full
3D disk, K = 32 neighbour slots, physical coagulation + fragmentation kernel, operator horizon
h_op = 0.05 per step. A step is one operator call evolve_local_collisions(h_op) over the whole swarm
with state resident in HBM; the geometry epoch (sites, sort, hierarchy, top-K) is built once before
the timed region, as in collision-only runs, and timed separately.

value = sum over timed steps of active owners / wall time (collision updates/s), one GPU, for the learned
operator (a majorant trained on samples of ten plain steps of the same swarm, Poisson thinning with exact
FP64 acceptance); the plain operator on the same swarm is timed first and reported beside it, with the
batched ensemble (64 realizations, plain and learned) and the clumped batched stress case (64
realizations totalling 1e7, vortex and ring overdensities) as secondary legs.
The timed steps run without per-kernel events; each timed leg is followed by --profile-steps calls of
the same workload with HIP events around every kernel launch (on the library's launch stream), which
give the per-kernel times. The roofline object is for the kernel with the largest measured time: its
algorithmic bytes per launch (owners x bytes per owner, listed in KERNEL_BYTES) over its mean HIP-event
duration.
The CPU baseline runs the oracle (host restatement of the same operator, OpenMP) on a bounded sample.

    python bench.py [--gpus 1] [--steps 100] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "collision updates/s at 1e7 reps, K=32: learned+batched vs plain, same GPU"
HBM_PEAK_GBS = 8000.0
# vector-ALU issue peak of MI355X: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles each (FP64 FMA
# included, at full rate), 2.4 GHz
VALU_PEAK_WAVE_INSTR = 256*4*2.4e9/4
K = 32
# algorithmic bytes (K = 32) of each timed kernel, and what limits it (measured: SQ counters and traces)
#   publish:      per due owner: state (s, N) read 16 + reservoir publish 16
#   mark_stale:   per due owner: its stale flag 1 (a lower bound: the reverse rows read are not counted)
#   bath_rate:    per exactly rated owner: codes 4K + partner (s, N) 16K + measure 8 + owner (y, z) 16 + group/active 5
#   screen:       per due owner: cached rate 8 + measure 8 + draw counter read + write 16 + particle index 4
#                 + published (s, N) 16
#   candidates:   per thinning candidate: candidate record 16 + owner state 16 + measure 8; per pair term of its
#                 K-sums: code 4 + partner (s, N) 16
#   queue_prefix: one block per phase over the group counters (not counted)
#   chain:        per chained owner 140 (state, clock, draw counter, sums), per pair term: code 4 + partner (s, N) 16
#   chain_audit:  per chained owner: queue entry 4 + chain sums 36 + published and final state 24
#   majorant:     per due owner: stale flag 1; per evaluated owner: codes 4K + FP32 partner records 8K + owner
#                 state 16 + measure 8 + environment 40 + bound 8
# concurrent streams (and host threads) of the ensemble legs: one per hardware queue of the process
ENSEMBLE_STREAMS = 4
PER_OWNER = {"publish": 32, "mark_stale": 1, "screen": 52}
RATED_BYTES = 20*K + 29
CAND_BYTES, CAND_PAIR_BYTES = 40, 20
CHAIN_OWNER_BYTES, CHAIN_PAIR_BYTES, AUDIT_OWNER_BYTES = 140, 20, 64
MAJORANT_DUE_BYTES, MAJORANT_EVAL_BYTES = 1, 12*K + 72
LIMIT = {"publish": "hbm", "mark_stale": "fp64-valu + gather latency", "bath_rate": "fp64-valu",
         "screen": "valu (draws) + latency", "candidates": "fp64-valu", "queue_prefix": "latency",
         "chain": "fp64-valu + serial event latency", "chain_audit": "latency",
         "majorant": "valu (features) + gather latency"}


def kernel_bytes(kernel: str, rec: dict, tot: dict) -> float:
    """Algorithmic bytes moved by all timed launches of one kernel, from the same calls' counters."""
    if kernel in PER_OWNER:
        return rec["owners"]*PER_OWNER[kernel]
    if kernel == "bath_rate":
        return tot["rated_owners"]*RATED_BYTES
    if kernel == "candidates":
        return tot["candidates"]*CAND_BYTES + tot.get("candidate_pairs", 0)*CAND_PAIR_BYTES
    if kernel == "majorant":
        return rec["owners"]*MAJORANT_DUE_BYTES + tot.get("bound_evals", 0)*MAJORANT_EVAL_BYTES
    if kernel == "chain":
        chain_pairs = tot["pair_evals"] - K*tot["rated_owners"] - tot.get("candidate_pairs", 0)
        return tot["chain_owners"]*CHAIN_OWNER_BYTES + chain_pairs*CHAIN_PAIR_BYTES
    if kernel == "chain_audit":
        return tot["chain_owners"]*AUDIT_OWNER_BYTES
    return 0.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> dict:
    """The host share this process runs on: nproc, the affinity set, OMP_NUM_THREADS and the CPU model."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), model)
    except OSError:
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = int(omp) if omp and omp.isdigit() and int(omp) > 0 else affinity
    return {"threads": threads, "nproc": os.cpu_count(), "affinity": affinity, "omp_num_threads": omp,
            "cpu_model": model}


def cpu_baseline(n_sample: int, steps: int, horizon: float, threads: int) -> dict:
    """The oracle's operator on host cores over a bounded sample of the same workload."""
    import numpy as np
    from gamedev_amd import synth
    from oracle.pyoracle import Oracle
    sys.path.insert(0, str(ROOT / "tests"))
    from helpers import root_cube

    o = Oracle("fiducial_3d")
    st = synth.fiducial_3d_state(n_sample, seed=0)
    t0 = time.perf_counter()
    pts, _, q, act, _ = o.site_init(st["x"], st["y"], st["z"], st["size"], st["numr"])
    org, w = root_cube(o.info)
    keys = o.morton_keys(pts, org, w, o.info["morton_max_level"], o.info["dim"])
    perm = o.sort_perm(keys)
    H = o.hierarchy(keys[perm], org, w, o.info["dim"], o.info["morton_leaf_target"], o.info["morton_max_level"])
    codes, d2 = o.topk_tree(pts, act, perm, H, o.info["dim"], o.info["morton_max_level"],
                            np.arange(n_sample, dtype=np.int32), K, q, threads=threads)
    codes[act == 0] = -1
    maxd2 = np.where(codes >= 0, d2, np.float32(0)).max(axis=1)
    meas = np.array([o.ball_measure(float(st["y"][i]), float(st["z"][i]), math.sqrt(float(maxd2[i])))
                     if act[i] else 0.0 for i in range(n_sample)])
    geo_s = time.perf_counter() - t0
    sp = o.space_bin(st["x"], st["y"], st["z"])
    size, numr, rng = st["size"].copy(), st["numr"].copy(), np.zeros(n_sample, np.uint64)
    ctl = o.new_ctl()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.operator(st["y"], st["z"], size, numr, act, sp, codes, meas, rng, 0, horizon, st["total_mass"], ctl,
                   threads=threads)
    op_s = time.perf_counter() - t0
    n_act = int(act.sum())
    return {"value": steps*n_act/op_s, "unit": "collision updates/s", "cores": threads, "kind": "port",
            "sample": f"{n_sample} representatives of the same 3D disk distribution, {steps} operator steps of "
                      f"h_op={horizon}, operator-only (geometry epoch {geo_s:.2f} s excluded)"}


def batch_leg(model_dir: str, count: int, steps: int, dev, samples: int, epochs: int) -> dict:
    """Ensemble of independent realizations of log-uniform size (floor(2^U[14,20]) representatives,
    seeds 1000..), each with its own lambda_0, controller and random key. Timed three ways over the same
    initial states: one batched call per step (shared launches, one stream), the ensemble split over four
    concurrent streams (gdx_col_operator_ensemble), and one call per realization at a time; then the
    four-stream ensemble with a majorant trained on the ensemble's own states. model_dir ensemble_2d is
    the C1 distribution with K = 32 (constant kernel: rates grow with the realization's size; the horizon
    gives the largest realization 5% events per step); fiducial_3d the 3D physical disk (rates fall with
    size: the smallest realizations carry the events)."""
    import numpy as np
    import torch
    from gamedev_amd import CollisionOperator, Majorant, evolve_batch, evolve_ensemble, synth

    sizes = synth.ensemble_sizes(count, 14, 20, seed=12345)
    if model_dir == "ensemble_2d":
        make_state = lambda r, n: synth.plain_2d_state(n, seed=1000 + r)
    else:
        make_state = lambda r, n: synth.fiducial_3d_state(n, seed=1000 + r)

    def build():
        ops = []
        for r, n in enumerate(sizes):
            st = make_state(r, n)
            d = synth.to_device(st, dev)
            op = CollisionOperator(model_dir, n, st["total_mass"], seed=r, defer_writeback=True)
            op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
            op.geometry()
            ops.append((op, d))
        return ops

    ops = build()
    if model_dir == "ensemble_2d":
        big = ops[int(np.argmax(sizes))][0]
        big.rate_pass()
        _, act, _ = big.sites()
        horizon = 0.05/float(big.cached()[act == 1, 0].mean())
    else:
        horizon = 0.05
    active = sum(op.active_count() for op, _ in ops)

    def timed(ops, call):
        """call(operators, steps) -> stats per step; one warmup step, then `steps` timed"""
        call([op for op, _ in ops], 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        per_step = call([op for op, _ in ops], steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        waves = np.array([[s["waves"] for s in out] for out in per_step])
        return steps*active/el, waves, [s for out in per_step for s in out]

    batch_steps = lambda o, k: [evolve_batch(o, horizon) for _ in range(k)]
    stream_steps = lambda o, k: evolve_ensemble(o, horizon, k, ENSEMBLE_STREAMS)
    one_at_a_time = lambda o, k: [[op.evolve(horizon) for op in o] for _ in range(k)]

    log(f"batch leg ({model_dir}): {count} realizations, {sum(sizes)} representatives, horizon {horizon:.4g}")
    batched, waves, _ = timed(ops, batch_steps)
    del ops
    ops = build()
    streams, _, st_e = timed(ops, stream_steps)
    # the same realizations advanced one operator call at a time
    sequential, _, _ = timed(ops, one_at_a_time)
    log(f"batch leg ({model_dir}): {batched:.3e} updates/s batched, {streams:.3e} on {ENSEMBLE_STREAMS} streams, "
        f"{sequential:.3e} one call at a time")
    out = {"model": model_dir, "realizations": count, "representatives": sum(sizes), "active_owners": active,
           "steps": steps, "horizon": horizon, "updates_per_s_batched": batched,
           "updates_per_s_ensemble_streams": streams, "ensemble_streams": ENSEMBLE_STREAMS,
           "updates_per_s_one_call_at_a_time": sequential, "batched_over_sequential": batched/sequential,
           "events_per_step": sum(s["events"] for s in st_e)/steps,
           "waves_per_step_mean": float(waves.mean()), "waves_per_step_max": int(waves.max())}
    # a majorant trained on the four largest realizations' states (two plain steps each)
    order = np.argsort(sizes)[::-1][:4]
    X, T = [], []
    per = max(1, samples//8)
    for _ in range(2):
        for r in order:
            f, S, R = ops[r][0].majorant_sample()
            ok = torch.nonzero((S > 0) & (R > 0)).squeeze(1)
            pick = ok[torch.randperm(len(ok), device=dev)[:per]]
            X.append(f[pick].clone())
            T.append(torch.log(R[pick]/S[pick]).float())
        evolve_ensemble([ops[r][0] for r in order], horizon, 1, ENSEMBLE_STREAMS)
    model = Majorant(model_dir, seed=1)
    losses, safety = model.fit(torch.cat(X), torch.cat(T), epochs=epochs, batch=8192, lr=1e-3, seed=1)
    del ops, X, T
    ops = build()
    for op, _ in ops:
        op.use_majorant(model, safety)
    learned, lwaves, stats = timed(ops, stream_steps)
    chk = sum(s["exact_checks"] for s in stats)
    cand = sum(s["candidates"] for s in stats)
    out.update({"updates_per_s_ensemble_learned": learned, "learned_over_plain_ensemble": learned/streams,
                "learned_realizations": int(sum(1 for op, _ in ops if op.active_count() >= 65536)),
                "learned_safety": safety, "learned_waves_per_step_mean": float(lwaves.mean()),
                "learned_candidates": int(cand),
                "learned_violation_rate": sum(s["violations"] for s in stats)/max(1, chk),
                "learned_acceptance_rate": sum(s["accepted"] for s in stats)/max(1, cand)})
    log(f"batch leg ({model_dir}): {learned:.3e} updates/s on {ENSEMBLE_STREAMS} streams with the majorant")
    del ops
    torch.cuda.empty_cache()
    return out


def ensemble_moments(ds) -> "np.ndarray":
    """M0..M3 of the mass distribution, sum_i N_i m_i^k, over every realization."""
    import numpy as np
    from gamedev_amd import synth
    M = np.zeros(4)
    for d in ds:
        s = d["size"].double().cpu().numpy()
        N = d["numr"].double().cpu().numpy()
        m = np.pi*synth.RHO_0*s**3/6.0
        M += [N.sum(), (N*m).sum(), (N*m*m).sum(), (N*m**3).sum()]
    return M


def wedge_leg(n: int, steps: int, dev) -> dict:
    """Azimuthal wedge model (x in [-0.1, 0.1], physical kernel, K=32): geometry epoch with periodic seam
    ghosts for the Morton search and with 3n image nodes for the k-d search, then plain operator steps
    at a horizon of a quarter of the controller's refresh bound."""
    import torch
    from gamedev_amd import CollisionOperator, synth
    st = synth.wedge_2d_state(n, seed=0)
    out = {"representatives": n, "steps": steps}
    for search in ("kdtree", "morton"):
        d = synth.to_device(st, dev)
        op = CollisionOperator("wedge_2d", n, st["total_mass"], seed=0, defer_writeback=True, search=search)
        op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
        op.geometry()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.geometry()
        torch.cuda.synchronize()
        out[f"geometry_epoch_ms_{search}"] = 1e3*(time.perf_counter() - t0)
        info = op.record_info()
        out[f"records_{search}"] = info["records"]
        out["unique_ids"] = info["unique_ids"]
    op.rate_pass()
    pts, act, _ = op.sites()
    c = op.cached()
    J2 = float(c[act == 1, 2].mean())
    h = 0.25*(op.info["col_bath_eps"]**2)/J2
    op.evolve(h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev = 0
    for _ in range(steps):
        ev += op.evolve(h)["events"]
    op.store_state()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    active = op.active_count()
    out.update({"horizon": h, "active_owners": active, "updates_per_s": steps*active/el,
                "ms_per_step": 1e3*el/steps, "events_per_step": ev/steps})
    log(f"wedge leg: {n} representatives, {out['records_morton']} Morton records "
        f"({out['geometry_epoch_ms_morton']:.1f} ms epoch), {out['records_kdtree']} k-d nodes "
        f"({out['geometry_epoch_ms_kdtree']:.1f} ms), {out['updates_per_s']:.3e} updates/s")
    return out


def clumped_leg(count: int, total: int, steps: int, horizon: float, dev, samples: int, epochs: int) -> dict:
    """The clumped stress case, batched: `count` realizations totalling `total` representatives (sizes
    floor(2^U[16,20]) rescaled), vortex and ring overdensities, sigma_ln = 1.5. The majorant is trained on
    clumped states (10 plain steps of eight extra realizations), then the plain and the learned operators
    advance identical copies of the ensemble, split over four concurrent streams (gdx_col_operator_ensemble);
    reports both rates, the violation and acceptance rates and the ensemble moments' relative difference."""
    import numpy as np
    import torch
    from gamedev_amd import CollisionOperator, Majorant, evolve_batch, evolve_ensemble, synth

    sizes = synth.ensemble_sizes(count, 16, 20, seed=54321, total=total)

    def build(ns, seed0, rng_seed0):
        ops = []
        for r, n in enumerate(ns):
            st = synth.clumped_3d_state(n, seed=seed0 + r)
            d = synth.to_device(st, dev)
            op = CollisionOperator("fiducial_3d", n, st["total_mass"], seed=rng_seed0 + r, defer_writeback=True)
            op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
            op.geometry()
            ops.append((op, d))
        return ops

    log(f"clumped leg: {count} realizations, {sum(sizes)} representatives")
    # training set: ten plain steps of eight realizations of the same distribution (other seeds)
    train_ops = build(sizes[:8], 9000, 500)
    per = max(1, samples//(10*len(train_ops)))
    X, T = [], []
    for _ in range(10):
        for op, _ in train_ops:
            f, S, R = op.majorant_sample()
            ok = torch.nonzero((S > 0) & (R > 0)).squeeze(1)
            pick = ok[torch.randperm(len(ok), device=dev)[:per]]
            X.append(f[pick].clone())
            T.append(torch.log(R[pick]/S[pick]).float())
        evolve_batch([op for op, _ in train_ops], horizon)
    del train_ops
    X, T = torch.cat(X), torch.cat(T)
    model = Majorant("fiducial_3d", seed=1)
    losses, safety = model.fit(X, T, epochs=epochs, batch=8192, lr=1e-3, seed=1)
    log(f"clumped majorant: {len(X)} samples, loss {losses[0]:.4f} -> {losses[-1]:.4f}, safety {safety:.3f}")
    out = {"realizations": count, "representatives": sum(sizes), "steps": steps, "horizon": horizon,
           "training_samples": int(len(X)), "first_epoch_loss": float(losses[0]),
           "last_epoch_loss": float(losses[-1]), "safety": safety}
    del X, T
    M_start = None
    for mode in ("plain", "learned"):
        ops = build(sizes, 7000, 0)
        if M_start is None:
            M_start = ensemble_moments([d for _, d in ops])
        if mode == "learned":
            for op, _ in ops:
                op.use_majorant(model, safety)
        active = sum(op.active_count() for op, _ in ops)
        # one untimed step: the ensemble call balances its streams by each realization's last waves
        evolve_ensemble([op for op, _ in ops], horizon, 1, ENSEMBLE_STREAMS)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = []
        for step_stats in evolve_ensemble([op for op, _ in ops], horizon, steps, ENSEMBLE_STREAMS):
            stats.extend(step_stats)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        for op, _ in ops:
            op.store_state()
        torch.cuda.synchronize()
        out[f"updates_per_s_{mode}"] = steps*active/el
        out[f"ms_per_step_{mode}"] = 1e3*el/steps
        out[f"events_{mode}"] = int(sum(s["events"] for s in stats))
        out[f"moments_{mode}"] = ensemble_moments([d for _, d in ops]).tolist()
        if mode == "learned":
            chk = sum(s["exact_checks"] for s in stats)
            cand = sum(s["candidates"] for s in stats)
            out["violation_rate"] = sum(s["violations"] for s in stats)/max(1, chk)
            out["acceptance_rate"] = sum(s["accepted"] for s in stats)/max(1, cand)
            out["exact_pair_evals_per_update"] = sum(s["pair_evals"] for s in stats)/max(1, steps*active)
        log(f"clumped leg {mode}: {out[f'updates_per_s_{mode}']:.3e} updates/s")
        del ops
        torch.cuda.empty_cache()
    mp, ml = np.array(out["moments_plain"]), np.array(out["moments_learned"])
    out["moment_rel_diff_learned_vs_plain"] = (ml/mp - 1.0).tolist()
    out["moment_change_plain"] = (mp/M_start - 1.0).tolist()
    out["learned_over_plain"] = out["updates_per_s_learned"]/out["updates_per_s_plain"]
    return out


def train_majorant(op, horizon: float, steps: int, samples: int, epochs: int, dev):
    """Training set from `steps` plain operator steps of the benchmark swarm (`samples` random owners in
    all, features and ln(R / S) at each step's start), then the library's BF16-MFMA training kernels
    (Adam, lr 1e-3, batch 8192) and the calibrated safety factor."""
    import torch
    from gamedev_amd import Majorant
    per = max(1, samples//steps)
    X, T = [], []
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    t0 = time.perf_counter()
    for _ in range(steps):
        f, S, R = op.majorant_sample()
        ok = torch.nonzero((S > 0) & (R > 0)).squeeze(1)
        pick = ok[torch.randperm(len(ok), device=dev, generator=gen)[:per]]
        X.append(f[pick].clone())
        T.append(torch.log(R[pick]/S[pick]).float())
        del f, S, R, ok, pick
        op.evolve(horizon)
    X = torch.cat(X)
    T = torch.cat(T)
    torch.cuda.synchronize()
    sample_s = time.perf_counter() - t0
    m = Majorant("fiducial_3d", seed=0)
    t0 = time.perf_counter()
    losses, safety = m.fit(X, T, epochs=epochs, batch=8192, lr=1e-3, seed=0)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    info = {"samples": int(len(X)), "sample_steps": steps, "epochs": epochs, "batch": 8192, "lr": 1e-3,
            "first_epoch_loss": float(losses[0]), "last_epoch_loss": float(losses[-1]), "safety": safety,
            "sampling_s": sample_s, "training_s": train_s}
    log(f"majorant trained on {len(X)} samples: loss {losses[0]:.4f} -> {losses[-1]:.4f}, safety {safety:.3f}, "
        f"{train_s:.2f} s")
    return m, safety, info


def timed_steps(op, steps: int, horizon: float, label: str) -> tuple[float, dict, int]:
    """Run `steps` operator calls; returns wall seconds, summed stats and active owners per step."""
    import torch
    keys = ("owner_updates", "events", "waves", "baths", "pair_evals", "chain_owners", "thinned", "candidates",
            "accepted", "exact_checks", "violations", "rated_owners", "candidate_pairs", "bound_evals")
    tot = {k: 0 for k in keys}
    tot["longest_chain"] = 0
    updates = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        s = op.evolve(horizon)
        updates += s["active_owners"]
        tot["longest_chain"] = max(tot["longest_chain"], s["max_chain_events"])
        for k in keys:
            tot[k] += s[k]
        if (i + 1) % 20 == 0:
            log(f"{label} step {i + 1}/{steps}")
    op.store_state()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, tot, updates


def profiled_window(op, steps: int, horizon: float, label: str) -> tuple[dict, dict, int]:
    """Kernel durations: `steps` further operator calls of the same workload with HIP events around every
    kernel launch. The timed steps run without them (the events cost about 0.2 ms per step)."""
    op.set_profile(True)
    op.reset_kernel_times()
    _, tot, _ = timed_steps(op, steps, horizon, f"{label} (kernel timing)")
    kt = op.kernel_times()
    op.set_profile(False)
    return kt, tot, steps


def kernel_table(kt: dict, tot: dict, steps: int) -> dict:
    """Per kernel: ms per step and per launch, algorithmic GB/s and its fraction of the HBM peak, the limiter."""
    out = {}
    for name, rec in kt.items():
        if rec["launches"] == 0:
            continue
        b = kernel_bytes(name, rec, tot)
        avg = rec["ms"]/rec["launches"]
        gbs = b/rec["launches"]/(avg*1e-3)/1e9 if avg > 0 else 0.0
        out[name] = {"ms_per_step": rec["ms"]/max(1, steps), "avg_launch_ms": avg, "launches": rec["launches"],
                     "algorithmic_GBps": gbs, "hbm_frac": gbs/HBM_PEAK_GBS, "limit": LIMIT.get(name, "")}
    return out


def roofline_of(kt: dict, tot: dict, dom: str | None = None, learned: bool = False) -> dict:
    dom = dom or max(kt, key=lambda k: kt[k]["ms"])
    rec = kt[dom]
    bytes_per_launch = kernel_bytes(dom, rec, tot)/max(1, rec["launches"])
    avg_s = rec["ms"]/max(1, rec["launches"])*1e-3
    achieved = bytes_per_launch/avg_s/1e9 if avg_s > 0 else 0.0
    traffic = None
    tfile = ROOT/"profiles"/f"traffic_{dom}_fiducial_3d.json"
    if tfile.exists():
        # HBM bytes per owner from the FETCH_SIZE / WRITE_SIZE counter runs, scaled to this run's launches
        per_owner = json.loads(tfile.read_text()).get("bytes_per_owner")
        if per_owner:
            traffic = per_owner*rec["owners"]/max(1, rec["launches"])
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved/HBM_PEAK_GBS, "traffic": traffic, "kernel": dom, "limit": LIMIT.get(dom, ""),
           "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_s*1e3, "launches": rec["launches"]}
    # learned-run kernels limited by the vector ALUs: their VALU instructions per launch (SQ_INSTS_VALU of the
    # committed counter run of the learned operator) over this run's launch time, against the issue peak
    cfile = ROOT/"profiles"/"counters_learned_kernels_fiducial_3d_1e7.json"
    if learned and "valu" in LIMIT.get(dom, "") and cfile.exists() and avg_s > 0:
        k = json.loads(cfile.read_text())["kernels"].get(dom)
        if k:
            rate = k["valu_instructions_per_launch"]/avg_s
            out["valu"] = {"achieved": rate, "peak": VALU_PEAK_WAVE_INSTR, "unit": "wave64 VALU instructions/s",
                           "frac": rate/VALU_PEAK_WAVE_INSTR, "busy_in_counter_run": k["valu_busy"],
                           "source": "profiles/counters_learned_kernels_fiducial_3d_1e7.json"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="operator calls with per-kernel HIP events after each timed leg (kernel times, roofline)")
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--horizon", type=float, default=0.05)
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--geometry-steps", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64,
                    help="realizations of the batched ensemble leg (0 skips it): sizes floor(2^U[14,20]), seeds 0..B-1")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--learned", type=int, default=1, help="1: train the majorant and time the learned operator")
    ap.add_argument("--train-steps", type=int, default=10)
    ap.add_argument("--train-samples", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--autotune", type=int, default=1, help="1: pick the Morton leaf target by timing the epoch")
    ap.add_argument("--clumped", type=int, default=64,
                    help="realizations of the clumped batched leg (0 skips it), totalling --clumped-total")
    ap.add_argument("--clumped-total", type=int, default=10_000_000)
    ap.add_argument("--clumped-steps", type=int, default=5)
    ap.add_argument("--wedge", type=int, default=1_000_000,
                    help="representatives of the azimuthal-wedge leg (0: skip)")
    ap.add_argument("--wedge-steps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first copy this many MiB device to device (a known-size dispatch that calibrates "
                         "hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the collision operator shards by independent realizations or "
              "Morton ranges (see DESIGN.md), and multi-GPU runs are not part of this build", file=sys.stderr)
        return 2

    import torch
    from gamedev_amd import CollisionOperator, synth

    dev = torch.device("cuda:0")
    if args.copy_mib > 0:
        a = torch.ones(args.copy_mib*(1 << 20)//8, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log(f"copied {args.copy_mib} MiB")
    log(f"generating {args.n} representatives")
    st = synth.fiducial_3d_state(args.n, seed=0)
    d = synth.to_device(st, dev)
    op = CollisionOperator("fiducial_3d", args.n, st["total_mass"], seed=0, defer_writeback=True)
    op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
    tune = None
    if args.autotune:
        tune = op.autotune((16, 32, 64, 128, 256), repeats=1)
        log(f"autotuned Morton leaf target {tune['best']}: " + ", ".join(f"{k}: {v:.1f} ms" for k, v in tune["ms"].items()))
        tile = op.autotune_tile((0, 8, 16, 32), repeats=1)
        log(f"autotuned top-K tile {tile['best']}: " + ", ".join(f"{k}: {v:.1f} ms" for k, v in tile["ms"].items()))
        block = op.autotune_block((1, 2, 4), repeats=1)
        log(f"autotuned top-K block {block['best']} wavefronts: "
            + ", ".join(f"{k}: {v:.1f} ms" for k, v in block["ms"].items()))
        tune = {"best": tune["best"], "ms": tune["ms"], "tile": tile, "block": block}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.geometry()
    torch.cuda.synchronize()
    geometry_ms = 1e3*(time.perf_counter() - t0)
    n_active = op.active_count()
    log(f"geometry epoch {geometry_ms:.1f} ms, {n_active} active owners")

    for i in range(args.warmup):
        op.evolve(args.horizon)
    log("warmup done")
    elapsed, tot, updates = timed_steps(op, args.steps, args.horizon, "plain")
    plain_value = updates/elapsed
    log(f"plain: {plain_value:.4e} updates/s, {1e3*elapsed/args.steps:.3f} ms/step")
    kt, p_tot, p_steps = profiled_window(op, args.profile_steps, args.horizon, "plain")

    # with a geometry rebuild before every operator call (transported runs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.geometry_steps):
        op.geometry()
        op.evolve(args.horizon)
    torch.cuda.synchronize()
    with_geo = args.geometry_steps*n_active/(time.perf_counter() - t0)
    # the transported Strang cadence (doc/guide_swarm.md:3013-3020): M steps of C(h/2) T(h) C(h/2) make
    # 2M operator calls of h/2 with M+1 geometry epochs and M-1 reused epochs (state screen only); the
    # transport itself is not run, its end of the epoch is the forced rebuild
    M = max(1, args.geometry_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.geometry()
    op.evolve(0.5*args.horizon)
    for k in range(M):
        op.geometry()
        op.evolve(0.5*args.horizon)
        if k + 1 < M:
            op.check_state()
            op.evolve(0.5*args.horizon)
    torch.cuda.synchronize()
    strang_s = time.perf_counter() - t0
    strang = {"transported_steps": M, "operator_calls": 2*M, "geometry_epochs": M + 1, "reused_epochs": M - 1,
              "half_horizon": 0.5*args.horizon, "wall_s": strang_s,
              "calls_per_s_x_owners": 2*M*n_active/strang_s, "transported_steps_per_s_x_owners": M*n_active/strang_s}
    log(f"Strang cadence: {M} steps, {2*M} half-calls, {M + 1} epochs in {1e3*strang_s:.1f} ms")

    learned = None
    model = None
    if args.learned:
        model, safety, train_info = train_majorant(op, args.horizon, args.train_steps, args.train_samples,
                                                   args.epochs, dev)
        op.use_majorant(model, safety)
        for i in range(args.warmup):
            op.evolve(args.horizon)
        l_elapsed, l_tot, l_updates = timed_steps(op, args.steps, args.horizon, "learned")
        learned_value = l_updates/l_elapsed
        log(f"learned: {learned_value:.4e} updates/s, {1e3*l_elapsed/args.steps:.3f} ms/step")
        l_kt, lp_tot, lp_steps = profiled_window(op, args.profile_steps, args.horizon, "learned")
        learned = {"value": learned_value, "ms_per_step": 1e3*l_elapsed/args.steps,
                   "learned_over_plain": learned_value/plain_value, "training": train_info,
                   "violation_rate": l_tot["violations"]/max(1, l_tot["exact_checks"]),
                   "acceptance_rate": l_tot["accepted"]/max(1, l_tot["candidates"]),
                   "candidates_per_owner_step": l_tot["candidates"]/max(1, l_updates),
                   "exact_pair_evals_per_update": l_tot["pair_evals"]/max(1, l_updates),
                   "events_per_s": l_tot["events"]/l_elapsed, "waves_per_step": l_tot["waves"]/args.steps,
                   "longest_chain_events": l_tot["longest_chain"],
                   "kernel_ms_per_step": {k: v["ms"]/max(1, lp_steps) for k, v in l_kt.items()},
                   "kernels": kernel_table(l_kt, lp_tot, lp_steps),
                   "roofline": roofline_of(l_kt, lp_tot, learned=True), "totals": l_tot}
        op.use_majorant(None)

    batched = batched_fid = None
    if args.batch > 0:
        del op, d
        torch.cuda.empty_cache()
        batched = batch_leg("ensemble_2d", args.batch, args.batch_steps, dev, args.train_samples, args.epochs)
        batched_fid = batch_leg("fiducial_3d", args.batch, args.batch_steps, dev, args.train_samples, args.epochs)

    clumped = None
    if args.clumped > 0:
        torch.cuda.empty_cache()
        clumped = clumped_leg(args.clumped, args.clumped_total, args.clumped_steps, args.horizon, dev,
                              args.train_samples, args.epochs)

    wedge = None
    if args.wedge > 0:
        torch.cuda.empty_cache()
        wedge = wedge_leg(args.wedge, args.wedge_steps, dev)

    roofline_plain = roofline_of(kt, p_tot)
    plain = {"value": plain_value, "ms_per_step": 1e3*elapsed/args.steps, "roofline": roofline_plain,
             "events_per_s": tot["events"]/elapsed, "pair_evals_per_s": tot["pair_evals"]/elapsed,
             "mean_baths_per_owner_step": tot["owner_updates"]/max(1, updates),
             "waves_per_step": tot["waves"]/args.steps, "chain_owners_per_step": tot["chain_owners"]/args.steps,
             "longest_chain_events": tot["longest_chain"],
             "kernel_ms_per_step": {k: v["ms"]/max(1, p_steps) for k, v in kt.items()},
             "kernels": kernel_table(kt, p_tot, p_steps)}
    workload = (f"fiducial 3D disk, single swarm of {args.n} representatives, K={K}, coagulation+fragmentation, "
                f"h_op={args.horizon}, operator-only with the state resident in HBM")
    if learned is not None:
        # the metric's subject: the learned operator on the 1e7-representative swarm, beside the plain one
        value, ms = learned["value"], learned["ms_per_step"]
        roofline = learned["roofline"]
        variant = ("learned: trained BF16 majorant, Poisson thinning with exact FP64 acceptance, exact 1/32 "
                   "subsample for the controller")
    else:
        value, ms, roofline, variant = plain_value, 1e3*elapsed/args.steps, roofline_plain, "plain"

    cpu = None
    if not args.no_cpu:
        host = host_cores()
        log(f"cpu baseline on {host['threads']} threads ({host['cpu_model']}, nproc {host['nproc']}, "
            f"affinity {host['affinity']}, OMP_NUM_THREADS {host['omp_num_threads']})")
        cpu = cpu_baseline(args.cpu_sample, args.cpu_steps, args.horizon, host["threads"])
        cpu["host"] = host

    out = {
        "metric": METRIC, "value": value, "unit": "collision updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{variant.split(':')[0]} operator, {workload}", "n": args.n, "k": K,
                   "horizon": args.horizon, "variant": variant},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": {
            "plain": plain,
            "learned": learned,
            "learned_over_plain": learned["learned_over_plain"] if learned else None,
            "geometry_epoch_ms": geometry_ms,
            "geometry_autotune": tune,
            "updates_per_s_with_geometry_rebuild": with_geo,
            "strang_cadence": strang,
            "active_owners": n_active,
            "batched_ensemble": batched,
            "batched_fiducial": batched_fid,
            "clumped_batched": clumped,
            "wedge": wedge,
        },
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
