#!/usr/bin/env python3
"""Collision-operator benchmark on one MI355X.

Workload (the fiducial model directory gamedev_amd/mod/fiducial_3d): 1e7 representatives in a full
3D disk, K = 32 neighbour slots, physical coagulation + fragmentation kernel, operator horizon
h_op = 0.05 per step. A step is one operator call evolve_local_collisions(h_op) over the whole swarm
with state resident in HBM; the geometry epoch (sites, sort, hierarchy, top-K) is built once before
the timed region, as in collision-only runs, and timed separately.

value = sum over timed steps of active owners / wall time (collision updates/s), one GPU.
The roofline object is for the kernel with the largest measured time: its algorithmic bytes per
launch (owners x bytes per owner, listed in KERNEL_BYTES) over its mean HIP-event duration.
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
K = 32
# algorithmic bytes (K = 32), per owner of the launch (all owners of the due groups) unless noted
#   publish:     owner (s, N) read 16 + reservoir publish 16
#   bath_rate:   codes 4K + partner (s, N) 16K + measure 8 + owner (y, z) 16 + group/active 5
#   screen:      cached rate 8 + measure 8 + RNG counter read + write 16 + particle index 4 + published (s, N) 16
#   chain:       per chained owner 140 (state, clock, RNG, sums), plus codes 4 + partner (s, N) 16 per pair term
#   chain_audit: per chained owner: queue entry 4 + chain sums 36 + published and final state 24
KERNEL_BYTES = {"publish": 32, "bath_rate": 20*K + 29, "screen": 52}
CHAIN_OWNER_BYTES, CHAIN_PAIR_BYTES, AUDIT_OWNER_BYTES = 140, 20, 64


def kernel_bytes(kernel: str, rec: dict, tot: dict) -> float:
    """Algorithmic bytes moved by all timed launches of one kernel."""
    if kernel in KERNEL_BYTES:
        return rec["owners"]*KERNEL_BYTES[kernel]
    chain_pairs = tot["pair_evals"] - K*tot["owner_updates"]
    if kernel == "chain":
        return tot["chain_owners"]*CHAIN_OWNER_BYTES + chain_pairs*CHAIN_PAIR_BYTES
    return tot["chain_owners"]*AUDIT_OWNER_BYTES


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def batch_leg(count: int, steps: int, horizon: float, dev) -> dict:
    """Ensemble of independent realizations of log-uniform size (floor(2^U[14,20]) representatives,
    seeds 0..count-1), each with its own lambda_0, controller and random key: one batched operator call
    per step (shared launches) against the same realizations advanced one call at a time."""
    import numpy as np
    import torch
    from gamedev_amd import CollisionOperator, evolve_batch, synth

    rng = np.random.default_rng(12345)
    sizes = [int(2.0**rng.uniform(14, 20)) for _ in range(count)]

    def build():
        ops = []
        for r, n in enumerate(sizes):
            st = synth.fiducial_3d_state(n, seed=1000 + r)
            d = synth.to_device(st, dev)
            op = CollisionOperator("fiducial_3d", n, st["total_mass"], seed=r, defer_writeback=True)
            op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
            op.geometry()
            ops.append((op, d))
        return ops

    log(f"batch leg: {count} realizations, {sum(sizes)} representatives")
    ops = build()
    active = sum(op.active_count() for op, _ in ops)
    evolve_batch([op for op, _ in ops], horizon)                  # warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    waves = []
    for _ in range(steps):
        waves.append([s["waves"] for s in evolve_batch([op for op, _ in ops], horizon)])
    torch.cuda.synchronize()
    batched = steps*active/(time.perf_counter() - t0)
    waves = np.array(waves)
    # the same realizations advanced one operator call at a time
    for op, _ in ops:
        op.evolve(horizon)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for op, _ in ops:
            op.evolve(horizon)
    torch.cuda.synchronize()
    sequential = steps*active/(time.perf_counter() - t0)
    log(f"batch leg: {batched:.3e} updates/s batched, {sequential:.3e} one call at a time")
    return {"realizations": count, "representatives": sum(sizes), "active_owners": active, "steps": steps,
            "updates_per_s_batched": batched, "updates_per_s_one_call_at_a_time": sequential,
            "batched_over_sequential": batched/sequential,
            "waves_per_step_mean": float(waves.mean()), "waves_per_step_max": int(waves.max())}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--horizon", type=float, default=0.05)
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--geometry-steps", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64,
                    help="realizations of the batched ensemble leg (0 skips it): sizes floor(2^U[14,20]), seeds 0..B-1")
    ap.add_argument("--batch-steps", type=int, default=10)
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
    op = CollisionOperator("fiducial_3d", args.n, st["total_mass"], seed=0, defer_writeback=True, profile=True)
    op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
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
    op.reset_kernel_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tot = {"owner_updates": 0, "events": 0, "waves": 0, "baths": 0, "pair_evals": 0, "chain_owners": 0}
    updates = 0
    longest_chain = 0
    for i in range(args.steps):
        s = op.evolve(args.horizon)
        updates += s["active_owners"]
        longest_chain = max(longest_chain, s["max_chain_events"])
        for k in tot:
            tot[k] += s[k]
        if (i + 1) % 20 == 0:
            log(f"step {i + 1}/{args.steps}")
    op.store_state()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = updates/elapsed
    kt = op.kernel_times()

    # with a geometry rebuild before every operator call (transported runs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.geometry_steps):
        op.geometry()
        op.evolve(args.horizon)
    torch.cuda.synchronize()
    with_geo = args.geometry_steps*n_active/(time.perf_counter() - t0)

    batched = None
    if args.batch > 0:
        del op, d
        torch.cuda.empty_cache()
        batched = batch_leg(args.batch, args.batch_steps, args.horizon, dev)

    dom = max(kt, key=lambda k: kt[k]["ms"])
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
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved/HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_s*1e3, "launches": rec["launches"]}

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(args.cpu_sample, args.cpu_steps, args.horizon, threads)

    out = {
        "metric": METRIC, "value": value, "unit": "collision updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3*elapsed/args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain operator (no learned majorant, single swarm): fiducial 3D disk, "
                               f"{args.n} representatives, K={K}, coagulation+fragmentation, h_op={args.horizon}, "
                               "operator-only with the state resident in HBM",
                   "n": args.n, "k": K, "horizon": args.horizon, "variant": "plain"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": {
            "geometry_epoch_ms": geometry_ms,
            "updates_per_s_with_geometry_rebuild": with_geo,
            "events_per_s": tot["events"]/elapsed,
            "pair_evals_per_s": tot["pair_evals"]/elapsed,
            "mean_baths_per_owner_step": tot["owner_updates"]/max(1, updates),
            "waves_per_step": tot["waves"]/args.steps,
            "chain_owners_per_step": tot["chain_owners"]/args.steps,
            "longest_chain_events": longest_chain,
            "kernel_ms_per_step": {k: v["ms"]/args.steps for k, v in kt.items()},
            "active_owners": n_active,
            "batched_ensemble": batched,
        },
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
