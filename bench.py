#!/usr/bin/env python3
"""Benchmark of the GeauxDock Monte Carlo hot path on one MI355X.

Workload (the plain one-job-per-launch configuration): one synthetic complex,
protein 3,000 effective points x 11 conformers, ligand 40 atoms x 20
conformers, 8 temperatures on a geometric ladder 0.004..0.04 -> 1,760
replicas; Gaussian moves N(0, 0.2 A) / N(0, 3 deg); temperature exchange every
100 MC steps.  One bench "step" = one exchange interval (100 MC steps of all
1,760 replicas + one exchange round), so the default --steps 100 is the whole
10,000-step job.  FP32 arithmetic, synthetic data (seeded generator).

Prints one JSON line (value = replica-steps/s).  The roofline object covers
the dominant kernel (gd_pair_kernel<8>): algorithmic FP32 flops per launch /
its mean launch time from HIP events recorded around every launch on the
job's stream.  cpu_baseline times the CPU oracle (a restatement of the
reference CPU backend) on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PRT_FLOPS, KDE_FLOPS, MCS_FLOPS = 62, 11, 9   # algorithmic flops per pair / match / entry
P, L, NLIG, NPRT, NTMP = 3000, 40, 20, 11, 8
STEPS_PER_EXCHANGE = 100
SEED = 2


def algorithmic_flops_per_step(job):
    """FP32 flops of one pair-kernel launch (one MC step of every replica)."""
    a = job.arrays()
    R = job.n_rep
    prt = R * job.desc.lig_natom * job.desc.prt_npoint
    kt, lt = a["kde_t"], a["lig_t"]
    kde = R * sum(int((kt == t).sum()) for t in lt)
    mcs = R * int((a["mcs_x"] != np.float32(0.01234)).sum())
    return PRT_FLOPS * prt + KDE_FLOPS * kde + MCS_FLOPS * mcs, prt


def cpu_baseline(job, mc_template, budget_s=15.0):
    """Time the CPU oracle on all replicas for a bounded number of MC steps."""
    from tests.oracle_ffi import Oracle
    import geauxdock_amd as g
    orc = Oracle()
    threads = min(16, os.cpu_count() or 1)
    # calibrate on one step, then run a sample of about budget_s seconds
    mc1 = g.make_mc(1, move_mode="gaussian", seed=SEED)
    t0 = time.perf_counter()
    orc.dock(job, mc1, threads=threads)
    t1 = time.perf_counter() - t0          # initial pass + 1 step
    per_step = max(t1 / 2.0, 1e-3)
    steps = max(1, min(200, int(budget_s / per_step)))
    mc = g.make_mc(steps, move_mode="gaussian", seed=SEED)
    _, _, ms = orc.dock(job, mc, threads=threads)
    # ms covers the initial pass + `steps` steps
    rs = job.n_rep * (steps + 1) / (ms / 1000.0)
    return {"value": rs, "unit": "replica-steps/s", "cores": threads, "kind": "port",
            "sample": "all %d replicas x %d MC steps (+ initial pass) of the same complex, "
                      "OpenMP over replicas" % (job.n_rep, steps)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="exchange intervals (100 MC steps each)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(jobs shard across GPUs without communication, see DESIGN.md)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import geauxdock_amd as g
    from geauxdock_amd import abi
    g.check(g.lib().gd_init(0), "gd_init")
    name = C.create_string_buffer(128)
    ncu, clk, hbm = C.c_int(), C.c_int(), C.c_size_t()
    g.lib().gd_device_info(name, 128, C.byref(ncu), C.byref(clk), C.byref(hbm))

    job = g.synth_complex(P, L, NLIG, NPRT, NTMP, floor_temp=0.004, ceiling_temp=0.04, seed=SEED)
    dev = g.DeviceJob(job)
    flops_step, prt_pairs = algorithmic_flops_per_step(job)
    # warmup: the same job for W exchange intervals (untimed)
    if args.warmup > 0:
        dev.dock(g.make_mc(args.warmup * STEPS_PER_EXCHANGE, steps_per_exchange=STEPS_PER_EXCHANGE,
                           move_mode="gaussian", seed=SEED), stats=False)
    steps = args.steps * STEPS_PER_EXCHANGE
    mc = g.make_mc(steps, steps_per_dump=steps, steps_per_exchange=STEPS_PER_EXCHANGE,
                   move_mode="gaussian", seed=SEED)
    t0 = time.perf_counter()
    out, st = dev.dock(mc, stats=True)     # returns after the device is idle
    wall = time.perf_counter() - t0
    rep_steps = job.n_rep * steps
    value = rep_steps / wall
    pair_ms = st.ms_pair_kernel / max(1, st.n_pair_launch)
    achieved = flops_step / (pair_ms / 1000.0) / 1e12
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pair_kernel_c2_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
    line = {
        "metric": "replica-steps/sec + docking jobs/sec, p50/p99 latency vs plain per-job launch",
        "value": value,
        "unit": "replica-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator: random-walk protein, random-tree ligand)",
        "config": {
            "workload": "plain one-job-per-launch: P=3000 points x 11 conformers, L=40 atoms x "
                        "20 conformers, 8 temperatures 0.004-0.04 (1760 replicas), 10000 MC steps, "
                        "exchange every 100",
            "replicas": job.n_rep, "mc_steps": steps, "prt_npoint": P, "lig_natom": L,
            "kde_npoint": job.desc.kde_npoint, "mcs_nrow": job.desc.mcs_nrow,
            "step_unit": "100 MC steps + 1 exchange round",
            "tiling": {"atoms_per_lane": dev.tiling()[0], "tile_points": dev.tiling()[1]},
            "device": name.value.decode(), "cu": ncu.value,
        },
        "jobs_per_s": 1.0 / wall * (10000.0 / steps),
        "pair_evals_per_s": prt_pairs * (steps + 1) / wall,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "kernel": "gd_pair_kernel<8>",
            "flops_per_launch": flops_step,
            "prt_pairs_per_launch": prt_pairs,
            "mean_launch_ms": pair_ms,
            "launches_timed": st.n_pair_launch,
            "prt_only_tflops": PRT_FLOPS * prt_pairs / (pair_ms / 1000.0) / 1e12,
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(job, mc)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
