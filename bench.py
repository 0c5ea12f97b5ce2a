#!/usr/bin/env python3
"""Benchmark of the GeauxDock Monte Carlo hot path on one MI355X.

Workload (the plain one-job-per-launch configuration): one synthetic complex,
protein 3,000 effective points x 11 conformers, ligand 40 atoms x 20
conformers, 8 temperatures on a geometric ladder 0.004..0.04 -> 1,760
replicas; Gaussian moves N(0, 0.2 A) / N(0, 3 deg); replica exchange every
100 MC steps along all three axes (temperature labels, then pose swaps between
neighbour ligand conformers and neighbour protein conformers, one extra energy
pass per conformer axis).  One bench "step" = one exchange interval (100 MC
steps of all 1,760 replicas + one exchange point), so the default --steps 100 is the whole
10,000-step job.  FP32 arithmetic, synthetic data (seeded generator).

Prints one JSON line (value = replica-steps/s, on the library's default two
streams).  The roofline object covers the dominant kernel (gd_pair_kernel<LC>,
LC ligand atoms per lane) on a pipeline basis: algorithmic FP32 flops of one
MC step's pair work / the steady-state MC step period, from HIP events at the
first stream's pair launch of every 16th MC step (the whole step, tail
kernels and exchange points included, is charged to the pair kernel); the
half-launch mean, the step's pair span and an isolated one-stream launch
are reported beside it.  cpu_baseline times the CPU oracle (a restatement of
the reference CPU backend, in the reference's FP32 expression order) on a
bounded sample of the same workload.

The same line carries, unless --no-extra:
  reference_shape  the headline job in the reference's plain shape (dumps of
                   200 steps, a blocking copy of each dump's records);
  c1               the CPU reference configuration in full (GPU, oracle and
                   the reference CPU backend itself where oracle/_ref is built);
  c3               the batched screen in full (256 ligands x 440 replicas x
                   5,000 steps against one 4,000-point target) against the plain
                   path in both shapes on 32 randomly drawn jobs, with its own
                   roofline and CPU baseline;
  c1_batched       256 complexes of the CPU reference configuration's shape
                   (distinct seeds; 25 replicas, 3,000 steps each) batched,
                   against the plain path in both shapes on 16 of them: the
                   small jobs one launch cannot fill the GPU with.
--workload c1|c1b|c3|c4|c5 print one of these alone (c1b: c1_batched, c4:
heterogeneous batch, c5: Poisson serving at 50/200/800 arrivals/s).
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
AXES = ("temp", "lig", "prt")   # replica exchange across temperature x ligand x protein conformer
SEED = 2


C2_JOB_STEPS = 10000   # MC steps of one docking job of the headline workload


def c2_rates(n_rep, mc_steps, wall_s):
    """(replica-steps/s, docking jobs/s) of a timed run of mc_steps MC steps of
    the headline job; one job is the whole C2_JOB_STEPS-step run, so the timed
    run is mc_steps / C2_JOB_STEPS of a job."""
    return n_rep * mc_steps / wall_s, (mc_steps / float(C2_JOB_STEPS)) / wall_s


def c2_workload(mc_steps):
    return ("plain one-job-per-launch: P=3000 points x 11 conformers, L=40 atoms x "
            "20 conformers, 8 temperatures 0.004-0.04 (1760 replicas), exchange every 100 "
            "MC steps (temperature, ligand conformer, protein conformer); %d MC steps of "
            "the %d-step job timed" % (mc_steps, C2_JOB_STEPS))


def algorithmic_flops_per_step(job):
    """FP32 flops of one pair-kernel launch (one MC step of every replica)."""
    a = job.arrays()
    R = job.n_rep
    prt = R * job.desc.lig_natom * job.desc.prt_npoint
    kt, lt = a["kde_t"], a["lig_t"]
    kde = R * sum(int((kt == t).sum()) for t in lt)
    mcs = R * int((a["mcs_x"] != np.float32(0.01234)).sum())
    return PRT_FLOPS * prt + KDE_FLOPS * kde + MCS_FLOPS * mcs, prt


def cpu_model():
    """The host CPU as /proc/cpuinfo names it, with its family / model numbers
    (the name alone can be generic under virtualisation)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                info[k] = v.strip()
            if line.strip() == "" and info:
                break
    except OSError:
        pass
    name = info.get("model name", "unknown")
    if "cpu family" in info:
        name += " (family %s model %s stepping %s)" % (info["cpu family"], info.get("model", "?"),
                                                       info.get("stepping", "?"))
    return name


_HOST_CORES = None


def host_cores():
    """(logical CPUs this process may run on, physical cores among them,
    physical cores of the whole host), read once: an OpenMP run with
    OMP_PROC_BIND pins the calling thread, after which its affinity mask
    shows one CPU."""
    global _HOST_CORES
    if _HOST_CORES is None:
        _HOST_CORES = _read_host_cores()
    return _HOST_CORES


def _read_host_cores():
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(os.cpu_count() or 1))

    def core_of(c):
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            return (open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip())
        except OSError:
            return ("?", str(c))
    phys_allowed = len({core_of(c) for c in allowed})
    n_all = os.cpu_count() or len(allowed)
    phys_all = len({core_of(c) for c in range(n_all)})
    return len(allowed), phys_allowed, phys_all


def cpu_threads():
    """Host threads for the CPU leg: one per physical core this process may
    use, at most 16 (the CPU share of a one-GPU job; OMP_NUM_THREADS, when set,
    caps it), threads bound (OMP_PROC_BIND)."""
    os.environ.setdefault("OMP_PROC_BIND", "true")
    _, phys, _ = host_cores()
    n = phys
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if env > 0:
        n = min(n, env)
    return max(1, min(16, n))


def cpu_fields(threads):
    """The CPU leg's hardware description: threads used, cores, model."""
    logical, phys, phys_all = host_cores()
    return {"cores": threads, "cpu_model": cpu_model(), "affinity_logical_cpus": logical,
            "affinity_physical_cores": phys, "host_physical_cores": phys_all,
            "threads_note": "one OpenMP thread per physical core, at most 16 (a one-GPU job's CPU share)"}


def oracle_rate(job, steps, threads, rep_hi=0):
    """replica-steps/s of the CPU oracle in the reference's own FP32 expression
    order (OpenMP over replicas) for `steps` MC steps of replicas [0, rep_hi)
    (0: all); the timed region is the MC loop, initial pass included."""
    from tests.oracle_ffi import Oracle
    import geauxdock_amd as g
    orc = Oracle()
    mc = g.make_mc(steps, move_mode="gaussian", seed=SEED)
    _, _, ms = orc.dock(job, mc, threads=threads, rep_hi=rep_hi, literal=True)
    n = rep_hi if rep_hi > 0 else job.n_rep
    return n * (steps + 1) / (ms / 1000.0), ms


def cpu_baseline(job, budget_s=15.0, note=""):
    """The CPU oracle on all replicas of `job` for a bounded number of MC steps
    (about budget_s seconds, calibrated on one step)."""
    threads = cpu_threads()
    hw = cpu_fields(threads)   # before OpenMP binds this thread to one core
    _, ms1 = oracle_rate(job, 1, threads)          # initial pass + 1 step
    per_step = max(ms1 / 2000.0, 1e-3)
    steps = max(1, min(200, int(budget_s / per_step)))
    rs, _ = oracle_rate(job, steps, threads)
    out = {"value": rs, "unit": "replica-steps/s", "kind": "port",
           "sample": "all %d replicas x %d MC steps (+ initial pass) of the same complex, OpenMP over "
                     "replicas, the reference's FP32 expression order%s" % (job.n_rep, steps, note)}
    out.update(hw)
    return out


def timed_plain(host, mc, reference_shape):
    """One job, one job per launch sequence, timed from the host complex to its
    results: device-job creation (packing + H2D), the run and the release.
    reference_shape: records on, a blocking copy of every dump's records to the
    host after the dump (backend_gpu.cu:96-118), as the reference's plain path."""
    import geauxdock_amd as g
    t0 = time.perf_counter()
    d = g.DeviceJob(host)
    if reference_shape:
        d.dock_stream(mc, None, blocking=True, stats=False)
    else:
        d.dock(mc, stats=False)
    d.close()
    return time.perf_counter() - t0


def roofline(flops, ms, kernel, **extra):
    achieved = flops / (ms / 1000.0) / 1e12
    r = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
         "frac": achieved / PEAK_FP32_TFLOPS, "kernel": kernel, "flops_per_launch": flops,
         "mean_launch_ms": ms}
    r.update(extra)
    return r


def c1_extra():
    """The CPU reference configuration in full: P = 1,200 points x 5 conformers,
    L = 24 atoms x 5 conformers, 1 temperature (25 replicas), 3,000 MC steps,
    Gaussian moves N(0, 0.2 A) / N(0, 3 deg), seed 1: the plain GPU path (its
    own shape and the reference's), the CPU oracle and, where it was built,
    the reference CPU backend itself."""
    import geauxdock_amd as g
    threads = cpu_threads()
    hw = cpu_fields(threads)
    job = g.synth_complex(1200, 24, 5, 5, 1, floor_temp=0.044, ceiling_temp=0.044, seed=1)
    mc = g.make_mc(3000, steps_per_dump=3000, move_mode="gaussian", seed=1)
    mc_ref = g.make_mc(3000, steps_per_dump=200, move_mode="gaussian", seed=1)
    flops, pairs = algorithmic_flops_per_step(job)
    dev = g.DeviceJob(job)
    dev.dock(g.make_mc(100, move_mode="gaussian", seed=1), stats=False)   # warm-up
    t0 = time.perf_counter()
    out, st = dev.dock(mc, stats=True)
    tg = time.perf_counter() - t0
    lc = dev.tiling()[0]
    dev.close()
    t_ref = timed_plain(job, mc_ref, True)
    pair_ms = st.ms_pair_kernel / max(1, st.n_pair_launch)
    rs_gpu = job.n_rep * 3000 / tg
    rs_cpu, cpu_ms = oracle_rate(job, 3000, threads)
    line = {"workload": "C1 in full: P=1200 points x 5 conformers, L=24 atoms x 5 conformers, 1 "
                        "temperature (25 replicas), 3000 MC steps, plain GPU path",
            "value": rs_gpu, "unit": "replica-steps/s", "ms": tg * 1000.0,
            "jobs_per_s": 1.0 / tg,
            "reference_shape": {"jobs_per_s": 1.0 / t_ref, "replica_steps_per_s": job.n_rep * 3000 / t_ref,
                                "shape": "steps_per_dump 200, blocking record copy after each dump; "
                                         "timed from the host complex to results"},
            "roofline": roofline(flops, pair_ms, "gd_pair_kernel<%d>" % lc,
                                 launches_timed=st.n_pair_launch, prt_pairs_per_launch=pairs,
                                 note="25 replicas are one 64-lane block: the launch fills a "
                                      "few CUs of 256"),
            "acceptance_gpu": float(out.n_accept.mean() / 3000)}
    cpu = {"value": rs_cpu, "unit": "replica-steps/s", "kind": "port", "ms": cpu_ms,
           "sample": "the whole job (25 replicas x 3000 MC steps + initial pass), OpenMP over "
                     "replicas, the reference's FP32 expression order"}
    cpu.update(hw)
    # the reference CPU backend itself, compiled from its sources into oracle/_ref
    lib = os.path.join(ROOT, "oracle", "_ref", "libgdref.so")
    if os.path.exists(lib):
        from geauxdock_amd import abi
        R = C.CDLL(lib)
        R.ref_dock.argtypes = [C.POINTER(abi.gd_job_desc), C.c_int, C.c_int, C.POINTER(C.c_float),
                               C.POINTER(C.c_double), C.POINTER(C.c_float)]
        rms = C.c_double()
        scale = (C.c_float * 6)(*[float(x) for x in mc.move_scale])
        if R.ref_dock(C.byref(job.desc), 3000, 200, scale, C.byref(rms), None) == 0:
            cpu = {"value": job.n_rep * 3001 / (rms.value / 1000.0), "unit": "replica-steps/s",
                   "kind": "reference", "ms": rms.value,
                   "sample": "the whole job: backend_cpu_mic Dock compiled from the reference's "
                             "sources (oracle/_ref), uniform moves of the same scales, its own RNG",
                   "port_value": rs_cpu}
            cpu.update(hw)
    line["cpu_baseline"] = cpu
    return line


def _sizes(rng, n, kind):
    """Per-job sizes of the batched / served workloads (seeded)."""
    out = []
    for i in range(n):
        L = int(rng.integers(12, 81))
        if kind == "c1b":   # the CPU reference configuration's shape
            out.append(dict(P=1200, L=24, nlig=5, nprt=5, ntmp=1))
        elif kind == "c3":
            out.append(dict(P=4000, L=L, nlig=10, nprt=11, ntmp=4, prot=3))
        else:
            P = int(round(np.exp(rng.uniform(np.log(1500), np.log(9000)))))
            nlig = 5 if kind == "c4" else int(rng.integers(5, 21))
            out.append(dict(P=P, L=L, nlig=nlig, nprt=11, ntmp=4, prot=1000 + i // 32))
    return out


def _hosts(kind, sizes, seed0):
    import geauxdock_amd as g
    hosts = []
    for i, z in enumerate(sizes):
        if kind == "c1b":
            hosts.append(g.synth_complex(z["P"], z["L"], z["nlig"], z["nprt"], z["ntmp"], floor_temp=0.044,
                                         ceiling_temp=0.044, seed=seed0 + i))
        elif kind == "c5":
            hosts.append(g.synth_complex(z["P"], z["L"], z["nlig"], z["nprt"], z["ntmp"], floor_temp=0.004,
                                         ceiling_temp=0.04, seed=seed0 + i))
        else:   # C3: one target for every ligand; C4: 32 ligands per target
            hosts.append(g.synth_screen(z["P"], z["prot"], z["L"], seed0 + i, z["nlig"], z["nprt"],
                                        z["ntmp"], floor_temp=0.004, ceiling_temp=0.04))
        if (i + 1) % 256 == 0:
            print("[%s] %d complexes generated" % (kind, i + 1), file=sys.stderr, flush=True)
    return hosts


def batched(kind, n_jobs, mc_steps, batch_size=0, plain_sample=0, roofline_steps=128, cpu=True):
    """Batched form (one launch sequence, persistent work-queue pair kernel,
    graph-captured segments, tilings from the tuned table) against the plain
    one-job-per-launch path on the same complexes: jobs/s of each, both timed
    from the host complexes to the results.  The plain path runs in its own
    shape and in the reference's (per-dump blocking record copies), on every
    job or, with plain_sample > 0, on a seeded sample of that many jobs (jobs/s
    = sample / its summed time).  The roofline comes from a separate timed run
    of roofline_steps steps of the whole batch (HIP events on the persistent
    pair kernel's launches)."""
    import geauxdock_amd as g
    from geauxdock_amd import abi
    rng = np.random.default_rng(3 if kind == "c3" else 4)
    sizes = _sizes(rng, n_jobs, kind)
    hosts = _hosts(kind, sizes, 100)
    # the CPU reference configuration's shape has one temperature and no exchange
    spe, axes = (0, ()) if kind == "c1b" else (100, AXES)
    mcs = [g.make_mc(mc_steps, steps_per_exchange=spe, move_mode="gaussian", seed=5, job_id=i,
                     axes=axes) for i in range(n_jobs)]
    mca = (abi.gd_mc_param * n_jobs)(*mcs)
    outs = [g.host.ReplicaOut(h.n_rep) for h in hosts]
    oa = (abi.gd_replica_out * n_jobs)(*[o.c for o in outs])
    # the process's first device job pays the runtime's one-time set-up
    # (about 0.15 s): taken here, outside both legs
    g.DeviceJob(hosts[0]).close()
    bs = batch_size if batch_size > 0 else n_jobs
    t_setup = t_run = 0.0
    t0 = time.perf_counter()
    for lo in range(0, n_jobs, bs):
        hi = min(n_jobs, lo + bs)
        ta = time.perf_counter()
        devs = [g.DeviceJob(h) for h in hosts[lo:hi]]
        arr = (C.c_void_p * (hi - lo))(*[d.h.value for d in devs])
        tb_ = time.perf_counter()
        mc_lo = C.cast(C.addressof(mca) + lo * C.sizeof(abi.gd_mc_param), C.POINTER(abi.gd_mc_param))
        out_lo = C.cast(C.addressof(oa) + lo * C.sizeof(abi.gd_replica_out), C.POINTER(abi.gd_replica_out))
        g.check(g.lib().gd_batch_dock(arr, hi - lo, mc_lo, out_lo, None), "gd_batch_dock")
        tc = time.perf_counter()
        for d in devs:
            d.close()
        t_setup += tb_ - ta
        t_run += tc - tb_
        print("[%s] batched leg: %d jobs in %.1f s" % (kind, hi, time.perf_counter() - t0),
              file=sys.stderr, flush=True)
    tb = time.perf_counter() - t0
    rep_steps = sum(h.n_rep for h in hosts) * mc_steps
    pair_evals = sum(h.n_rep * h.desc.lig_natom * h.desc.prt_npoint for h in hosts) * (mc_steps + 1)
    # roofline: the persistent pair kernel of the first batch, event-timed
    nb = min(n_jobs, bs)
    devs = [g.DeviceJob(h) for h in hosts[:nb]]
    arr = (C.c_void_p * nb)(*[d.h.value for d in devs])
    rmc = (abi.gd_mc_param * nb)(*[g.make_mc(roofline_steps, steps_per_exchange=spe, seed=5, job_id=i,
                                             axes=axes) for i in range(nb)])
    st = abi.gd_stats()
    g.check(g.lib().gd_batch_dock(arr, nb, rmc, None, C.byref(st)), "gd_batch_dock (timed)")
    flops = sum(algorithmic_flops_per_step(h)[0] for h in hosts[:nb])
    pairs = sum(algorithmic_flops_per_step(h)[1] for h in hosts[:nb])
    tiling = sorted({d.tiling() for d in devs})
    for d in devs:
        d.close()
    pair_ms = st.ms_pair_kernel / max(1, st.n_pair_launch)
    # plain legs
    which = list(range(n_jobs))
    if 0 < plain_sample < n_jobs:
        which = sorted(np.random.default_rng(11).choice(n_jobs, plain_sample, replace=False).tolist())
    tp = tr = 0.0
    for q, i in enumerate(which):
        tp += timed_plain(hosts[i], mcs[i], False)
        mc_r = g.make_mc(mc_steps, steps_per_dump=200, steps_per_exchange=spe, move_mode="gaussian",
                         seed=5, job_id=i, axes=axes)
        tr += timed_plain(hosts[i], mc_r, True)
        if (q + 1) % 16 == 0:
            print("[%s] plain legs: %d jobs in %.1f s" % (kind, q + 1, tp + tr), file=sys.stderr, flush=True)
    plain_jps, ref_jps = len(which) / tp, len(which) / tr
    # each sampled job's bits against its own plain run at the shared tiling
    same = []
    for i in which[:2]:
        dp = g.DeviceJob(hosts[i])
        lc, pt, _ = g.host.tiling_lookup(hosts[i].desc.lig_natom, hosts[i].desc.prt_npoint,
                                         hosts[i].n_rep, True)
        dp.set_tiling(lc, pt)
        o, _ = dp.dock(mcs[i], stats=False)
        dp.close()
        same.append(bool(np.array_equal(o.best_energy, outs[i].best_energy)))
    line = {"workload": kind, "jobs": n_jobs, "mc_steps": mc_steps,
            "replicas": int(sum(h.n_rep for h in hosts)),
            "value": n_jobs / tb, "unit": "jobs/s",
            "replica_steps_per_s": rep_steps / tb, "pair_evals_per_s": pair_evals / tb,
            "timed": "host complexes to results (device-job creation, run, release), both legs",
            "batched_setup_s": t_setup, "batched_run_s": t_run, "batch_size": bs,
            "tilings": [list(t) for t in tiling],
            "plain_jobs_per_s": plain_jps, "plain_reference_shape_jobs_per_s": ref_jps,
            "plain_jobs_timed": len(which),
            "plain_sample": ("all jobs" if len(which) == n_jobs else
                             "%d of %d jobs drawn at random (seed 11), each timed alone" % (len(which), n_jobs)),
            "ratio_vs_plain": (n_jobs / tb) / plain_jps,
            "ratio_vs_plain_reference_shape": (n_jobs / tb) / ref_jps,
            "target_10x_met": (n_jobs / tb) / ref_jps >= 10.0,
            "best_energy_bitwise_equal_to_plain": same,
            "roofline": roofline(flops, pair_ms, "gd_pair_kernel_wq<4>", launches_timed=st.n_pair_launch,
                                 prt_pairs_per_launch=pairs,
                                 measured="%d MC steps of the first %d jobs, every 16th pair launch "
                                          "event-timed" % (roofline_steps, nb))}
    if cpu:
        rs, ms = oracle_rate(hosts[0], 20, cpu_threads())
        cb = {"value": rs, "unit": "replica-steps/s", "kind": "port",
              "jobs_per_s_extrapolated": rs / (hosts[0].n_rep * mc_steps) *
              (hosts[0].n_rep * hosts[0].desc.lig_natom / np.mean([h.n_rep * h.desc.lig_natom for h in hosts])),
              "sample": "job 0 (%d replicas, L=%d) for 20 MC steps + initial pass, OpenMP over replicas, "
                        "the reference's FP32 expression order; jobs/s scaled by the mean job's "
                        "replicas x ligand atoms" % (hosts[0].n_rep, hosts[0].desc.lig_natom)}
        cb.update(cpu_fields(cpu_threads()))
        line["cpu_baseline"] = cb
    return line


def fifo_latencies(arrivals, service):
    """Latency (arrival -> result) of each request at a server that runs one
    job at a time in arrival order, given each job's run time alone: a job
    starts when it has arrived and the previous job has finished."""
    lat, free = [], 0.0
    for a, t in zip(arrivals, service):
        free = max(free, a) + t
        lat.append(free - a)
    return np.array(lat)


def serve(rates, n_jobs, step_scale, max_active=32):
    """Serving: Poisson arrivals, admission by device footprint, continuous
    batching; p50/p99 latency (arrival -> result) and jobs/s per rate, and the
    plain path on the same requests (both shapes): jobs/s, and the p50/p99
    latency of a one-job-at-a-time server on the same arrivals, from each
    job's measured plain run time."""
    import geauxdock_amd as g
    from geauxdock_amd import abi
    res = []
    arrivals = {}
    rng = np.random.default_rng(5)
    sizes = _sizes(rng, n_jobs, "c5")
    steps = [int(rng.integers(2000, 10001) * step_scale) for _ in range(n_jobs)]
    hosts = _hosts("c5", sizes, 500)
    mcs = [g.make_mc(steps[i], steps_per_exchange=100, move_mode="gaussian", seed=7, job_id=i,
                     axes=AXES) for i in range(n_jobs)]
    for lam in rates:
        arr_t = np.cumsum(np.random.default_rng(int(lam)).exponential(1.0 / lam, n_jobs))
        arrivals[lam] = arr_t
        cfg = abi.gd_server_config(hbm_budget=200 << 30, max_active=max_active, sweep_steps=50)
        s = C.c_void_p()
        g.check(g.lib().gd_server_create(C.byref(cfg), C.byref(s)), "gd_server_create")
        t0 = time.perf_counter()
        sub, done, lat = 0, 0, []
        submit_t = {}
        while done < n_jobs:
            now = time.perf_counter() - t0
            while sub < n_jobs and arr_t[sub] <= now:
                t = C.c_int64()
                g.check(g.lib().gd_server_submit(s, C.byref(hosts[sub].desc), C.byref(mcs[sub]), None,
                                                 C.byref(t)), "submit")
                submit_t[t.value] = arr_t[sub]
                sub += 1
            na, nd = C.c_int(), C.c_int()
            g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "gd_server_step")
            rr = (abi.gd_job_result * 64)()
            n = C.c_int()
            g.lib().gd_server_results(s, rr, 64, C.byref(n))
            end = time.perf_counter() - t0
            for k in range(n.value):
                lat.append(end - submit_t[rr[k].ticket])
            if (done + n.value) // 50 > done // 50:
                print("[c5] rate %d/s: %d of %d jobs done in %.1f s" % (lam, done + n.value, n_jobs, end),
                      file=sys.stderr, flush=True)
            done += n.value
            if na.value == 0 and sub < n_jobs:
                time.sleep(max(0.0, arr_t[sub] - (time.perf_counter() - t0)))
        wall = time.perf_counter() - t0
        g.lib().gd_server_destroy(s)
        lat = np.array(lat)
        res.append({"rate_per_s": lam, "jobs": n_jobs, "jobs_per_s": n_jobs / wall,
                    "p50_latency_s": float(np.percentile(lat, 50)),
                    "p99_latency_s": float(np.percentile(lat, 99))})
    t_plain, t_ref = [], []
    for i in range(n_jobs):
        t_plain.append(timed_plain(hosts[i], mcs[i], False))
        mc_r = g.make_mc(steps[i], steps_per_dump=200, steps_per_exchange=100, move_mode="gaussian",
                         seed=7, job_id=i, axes=AXES)
        t_ref.append(timed_plain(hosts[i], mc_r, True))
        if (i + 1) % 50 == 0:
            print("[c5] plain: %d jobs in %.1f s" % (i + 1, sum(t_plain) + sum(t_ref)), file=sys.stderr,
                  flush=True)
    for r in res:
        for key, t in (("plain", t_plain), ("plain_reference_shape", t_ref)):
            lat = fifo_latencies(arrivals[r["rate_per_s"]], t)
            r[key + "_p50_latency_s"] = float(np.percentile(lat, 50))
            r[key + "_p99_latency_s"] = float(np.percentile(lat, 99))
    return {"workload": "c5-serving", "step_scale": step_scale, "max_active": max_active, "per_rate": res,
            "plain_jobs_per_s": n_jobs / sum(t_plain), "plain_reference_shape_jobs_per_s": n_jobs / sum(t_ref),
            "plain_note": "same requests, one job per launch sequence, back to back; plain latencies are "
                          "those of a one-job-at-a-time server on the same arrival trace, from each job's "
                          "measured run time alone"}


def _counter_rows(run_dir, counter):
    """(kernel name, grid size, value) of every dispatch's reading of one
    counter, from a rocprofv3 --pmc --output-format csv directory."""
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(run_dir, "*counter_collection.csv")):
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] == counter:
                rows.append((r["Kernel_Name"], int(r.get("Grid_Size") or 0), float(r["Counter_Value"])))
    return rows


def _counter_means(run_dir, counter):
    """Mean per-dispatch value of one counter, by kernel name."""
    vals = {}
    for k, _, v in _counter_rows(run_dir, counter):
        vals.setdefault(k, []).append(v)
    return {k: sum(v) / len(v) for k, v in vals.items()}


def traffic_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, out_json):
    """HBM bytes per pair-kernel launch from FETCH_SIZE / WRITE_SIZE runs (KiB per
    dispatch), each scaled by the same counter's reading for a plain 256 MiB copy."""
    n = 256 << 20
    cf = _counter_means(copy_fetch_dir, "FETCH_SIZE")
    cw = _counter_means(copy_write_dir, "WRITE_SIZE")
    copy_k = [k for k in cf if "copy" in k.lower() or "elementwise" in k]
    kf = max(copy_k, key=lambda k: cf[k])          # the copy: the largest reader
    fetch_scale = n / (cf[kf] * 1024.0)
    kw = kf if kf in cw else max(cw, key=lambda k: cw[k])
    write_scale = n / (cw[kw] * 1024.0)
    f = _counter_means(fetch_dir, "FETCH_SIZE")
    w = _counter_means(write_dir, "WRITE_SIZE")
    res = {"copy_bytes": n, "copy_kernel": kf, "fetch_scale": fetch_scale, "write_scale": write_scale,
           "kernels": {}}
    for k in f:
        res["kernels"][k] = {"fetch_kib": f[k], "write_kib": w.get(k, 0.0),
                             "hbm_bytes": f[k] * 1024.0 * fetch_scale + w.get(k, 0.0) * 1024.0 * write_scale}
    pair = [k for k in res["kernels"] if "gd_pair_kernel<" in k]
    if pair:
        # Two-stream steps launch the pair kernel as two half-grids; the whole
        # grid is the exchange points' swap passes and the one-stream launches.
        # Per MC step: the two halves' bytes.
        def by_grid(rows):
            d = {}
            for k, gs, v in rows:
                if k == pair[0]:
                    d.setdefault(gs, []).append(v)
            return {gs: sum(v) / len(v) for gs, v in d.items()}
        fg = by_grid(_counter_rows(fetch_dir, "FETCH_SIZE"))
        wg = by_grid(_counter_rows(write_dir, "WRITE_SIZE"))
        full = max(fg)
        kib = lambda gs: fg[gs] * 1024.0 * fetch_scale + wg.get(gs, 0.0) * 1024.0 * write_scale
        res["pair_by_grid"] = {str(gs): {"fetch_kib": fg[gs], "write_kib": wg.get(gs, 0.0),
                                         "hbm_bytes": kib(gs)} for gs in sorted(fg)}
        halves = [gs for gs in fg if gs != full]
        res["hbm_bytes_whole_launch"] = kib(full)
        if len(halves) == 1:
            res["hbm_bytes_per_launch"] = 2.0 * kib(halves[0])
        else:
            res["hbm_bytes_per_launch"] = sum(kib(gs) for gs in halves) if halves else kib(full)
        res["basis"] = ("one launch = one MC step's pair work: the sum of the two half-grid dispatches"
                        if halves else "one whole-grid dispatch")
        res["kernel"] = pair[0]
    json.dump(res, open(out_json, "w"), indent=1)
    return res


def trace_summary(trace_csv):
    """Per-kernel dispatch counts and mean durations from a rocprofv3
    kernel-trace CSV, and the pair kernel's per-step figure to set against
    bench.py's event-timed mean: for one-stream steps the dispatch duration;
    for two-stream steps (two half-dispatches on different streams, launched
    back to back) the span from the first start to the last end of each pair
    of halves.  Whole-job pair dispatches of the exchange points' swap passes
    are reported apart."""
    import csv
    rows = list(csv.DictReader(open(trace_csv)))
    kern = {}
    for r in rows:
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        k = kern.setdefault(r["Kernel_Name"], [0, 0.0])
        k[0] += 1
        k[1] += d
    out = {"kernels": {k: {"dispatches": v[0], "mean_us": v[1] / v[0] / 1000.0} for k, v in kern.items()}}
    pair = sorted((r for r in rows if "gd_pair_kernel<" in r["Kernel_Name"]),
                  key=lambda r: int(r["Correlation_Id"]))
    if pair:
        grids = sorted({int(r["Grid_Size_X"]) for r in pair})
        full = max(grids)
        halves = [r for r in pair if int(r["Grid_Size_X"]) != full]
        whole = [r for r in pair if int(r["Grid_Size_X"]) == full]
        spans = []
        i = 0
        while i + 1 < len(halves):
            a, b = halves[i], halves[i + 1]
            if a["Stream_Id"] != b["Stream_Id"] and int(a["Grid_Size_X"]) + int(b["Grid_Size_X"]) == full:
                spans.append(max(int(a["End_Timestamp"]), int(b["End_Timestamp"])) -
                             min(int(a["Start_Timestamp"]), int(b["Start_Timestamp"])))
                i += 2
            else:
                i += 1
        dur = lambda rs: sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in rs) / max(1, len(rs))
        # steady-state step period: successive starts of the first half's
        # dispatches (the stream of the first half-dispatch), gaps over 3x the
        # median (run and warm-up boundaries) left out
        period = None
        if halves:
            sa = halves[0]["Stream_Id"]
            starts = sorted(int(r["Start_Timestamp"]) for r in halves if r["Stream_Id"] == sa)
            d = [b - a for a, b in zip(starts, starts[1:])]
            if d:
                med = sorted(d)[len(d) // 2]
                keep = [x for x in d if x <= 3 * med]
                period = sum(keep) / len(keep) / 1000.0
        out["pair"] = {"whole_job_dispatches": len(whole), "whole_job_mean_us": dur(whole) / 1000.0,
                       "two_stream_steps": len(spans),
                       "two_stream_step_span_mean_us": (sum(spans) / len(spans) / 1000.0) if spans else None,
                       "half_dispatch_mean_us": dur(halves) / 1000.0 if halves else None,
                       "two_stream_step_period_us": period}
    return out


def copy_calibration(n_bytes=256 << 20, reps=3):
    """A plain device-to-device copy of known size (the library's own 16-byte
    vector copy kernel, gd_copy_kernel): the HBM counters (FETCH_SIZE,
    WRITE_SIZE) of its dispatches against n_bytes give the scale of those
    counters on this device."""
    import geauxdock_amd as g
    ms = C.c_double()
    g.check(g.lib().gd_device_copy(n_bytes, reps, C.byref(ms)), "gd_device_copy")
    return {"workload": "copy-calibration", "kernel": "gd_copy_kernel", "bytes_read_per_copy": n_bytes,
            "bytes_written_per_copy": n_bytes, "copies": reps, "ms_per_copy": ms.value,
            "gb_per_s": 2 * n_bytes / (ms.value / 1000.0) / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="exchange intervals (100 MC steps each)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", choices=["c1", "c1b", "c2", "c3", "c4", "c5", "copy"], default="c2",
                    help="c2 = the headline line; c1 the CPU reference configuration on CPU and GPU; "
                         "c3/c4 batched vs plain; c5 serving (extra lines); "
                         "copy = a plain 256 MiB device copy, the calibration for HBM counters")
    ap.add_argument("--traffic", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="summarise rocprofv3 FETCH_SIZE/WRITE_SIZE runs into "
                         "profiles/pair_kernel_c2_traffic.json (no GPU needed)")
    ap.add_argument("--trace-summary", metavar="KERNEL_TRACE_CSV",
                    help="summarise a rocprofv3 --kernel-trace CSV (no GPU needed)")
    ap.add_argument("--jobs", type=int, default=256, help="c3/c4/c5: jobs (requests)")
    ap.add_argument("--mc-steps", type=int, default=5000, help="c3/c4: MC steps per job; c5: "
                    "the step scale is mc_steps / 10000 of the 2,000-10,000 drawn per request")
    ap.add_argument("--batch-size", type=int, default=0,
                    help="c3/c4: jobs per batched launch sequence (0: all jobs in one)")
    ap.add_argument("--plain-sample", type=int, default=32,
                    help="c3/c4: plain legs on this many randomly drawn jobs (0: every job)")
    ap.add_argument("--max-active", type=int, default=32, help="c5: jobs the server keeps active at once")
    ap.add_argument("--no-extra", action="store_true",
                    help="default run: the headline line only, without its c1/c3 keys")
    args = ap.parse_args()
    host_cores()   # the CPU topology, before any OpenMP region binds this thread
    if args.trace_summary:
        print(json.dumps(trace_summary(args.trace_summary), indent=1))
        return
    if args.traffic:
        print(json.dumps(traffic_summary(*args.traffic, os.path.join(ROOT, "profiles",
                                                                      "pair_kernel_c2_traffic.json"))))
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported "
              "(jobs shard across GPUs without communication)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import geauxdock_amd as g
    from geauxdock_amd import abi
    g.check(g.lib().gd_init(0), "gd_init")
    if args.workload == "copy":
        print(json.dumps(copy_calibration()), flush=True)
        return
    if args.workload == "c1":
        print(json.dumps(c1_extra()), flush=True)
        return
    if args.workload in ("c1b", "c3", "c4"):
        print(json.dumps(batched(args.workload, args.jobs, args.mc_steps, args.batch_size,
                                 args.plain_sample, cpu=not args.no_cpu)), flush=True)
        return
    if args.workload == "c5":
        print(json.dumps(serve([50, 200, 800], args.jobs, args.mc_steps / 10000.0, args.max_active)), flush=True)
        return
    name = C.create_string_buffer(128)
    ncu, clk, hbm = C.c_int(), C.c_int(), C.c_size_t()
    g.lib().gd_device_info(name, 128, C.byref(ncu), C.byref(clk), C.byref(hbm))
    sclk = C.c_double()
    g.check(g.lib().gd_measure_sclk(200, C.byref(sclk)), "gd_measure_sclk")

    job = g.synth_complex(P, L, NLIG, NPRT, NTMP, floor_temp=0.004, ceiling_temp=0.04, seed=SEED)
    dev = g.DeviceJob(job)
    flops_step, prt_pairs = algorithmic_flops_per_step(job)
    # warmup: the same job for W exchange intervals (untimed)
    if args.warmup > 0:
        dev.dock(g.make_mc(args.warmup * STEPS_PER_EXCHANGE, steps_per_exchange=STEPS_PER_EXCHANGE,
                           move_mode="gaussian", seed=SEED, axes=AXES), stats=False)
    steps = args.steps * STEPS_PER_EXCHANGE
    mc = g.make_mc(steps, steps_per_dump=steps, steps_per_exchange=STEPS_PER_EXCHANGE,
                   move_mode="gaussian", seed=SEED, axes=AXES)
    t0 = time.perf_counter()
    out, st = dev.dock(mc, stats=True)     # returns after the device is idle
    wall = time.perf_counter() - t0
    # the engine clock under an all-CU FMA load, right after the timed region
    sclk_after = C.c_double()
    g.check(g.lib().gd_measure_sclk(200, C.byref(sclk_after)), "gd_measure_sclk")
    value, jobs_per_s = c2_rates(job.n_rep, steps, wall)
    # Two streams step the two halves of the replica blocks, so a step's two
    # pair half-launches overlap each other and the other half's reduce and
    # finish kernels: the pair kernel's rate is its flops per MC step over the
    # steady-state step period (HIP events at the pair launches of every 16th
    # step, on the pair kernel's stream), which charges the whole step -- tail
    # kernels and exchange points included -- to the pair kernel.
    period_ms = st.ms_step_period
    achieved = flops_step / (period_ms / 1000.0) / 1e12
    half_ms = st.ms_pair_kernel / max(1, st.n_pair_launch)
    # the same kernel alone: one stream, one launch per step, event-timed
    dev.set_streams(1)
    _, st1 = dev.dock(g.make_mc(400, steps_per_dump=400, steps_per_exchange=STEPS_PER_EXCHANGE,
                                move_mode="gaussian", seed=SEED, axes=AXES), stats=True)
    dev.set_streams(2)
    iso_ms = st1.ms_pair_kernel / max(1, st1.n_pair_launch)
    iso = flops_step / (iso_ms / 1000.0) / 1e12
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
            "workload": c2_workload(steps),
            "replicas": job.n_rep, "mc_steps": steps, "prt_npoint": P, "lig_natom": L,
            "kde_npoint": job.desc.kde_npoint, "mcs_nrow": job.desc.mcs_nrow,
            "step_unit": "100 MC steps + 1 exchange point",
            "exchange_axes": list(AXES),
            "swaps_accepted": {"temperature": int(out.n_exchange.sum()) // 2,
                               "ligand_conformer": int(out.n_swap_lig.sum()) // 2,
                               "protein_conformer": int(out.n_swap_prt.sum()) // 2},
            "tiling": {"atoms_per_lane": dev.tiling()[0], "tile_points": dev.tiling()[1]},
            "device": name.value.decode(), "cu": ncu.value, "hbm_bytes": hbm.value,
            "sclk_mhz_nominal": clk.value / 1000.0,
            "sclk_mhz_measured": [sclk.value, sclk_after.value],
            "sclk_note": "engine clock under an all-CU FMA load (shader-clock counter against the "
                         "constant-rate counter), before and after the timed region",
            "peak_from_device_tflops": ncu.value * 64 * 2 * 2 * clk.value * 1e3 / 1e12,
            "peak_from_measured_clock_tflops": ncu.value * 64 * 2 * 2 * sclk_after.value * 1e6 / 1e12,
        },
        "jobs_per_s": jobs_per_s,
        "pair_evals_per_s": prt_pairs * (steps + 1) / wall,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "kernel": "gd_pair_kernel<%d>" % dev.tiling()[0],
            "flops_per_launch": flops_step,
            "prt_pairs_per_launch": prt_pairs,
            "basis": "one launch = one MC step's pair work, two half-launches on two streams; its "
                     "duration = the steady-state MC step period (events at the pair launches of every "
                     "16th step on the pair kernel's stream), the whole step charged to the pair kernel",
            "mean_launch_ms": period_ms,
            "launches_timed": st.n_pair_launch // 2,
            "half_launch_mean_ms": half_ms,
            "pair_span_mean_ms": st.ms_pair_span,
            "prt_only_tflops": PRT_FLOPS * prt_pairs / (period_ms / 1000.0) / 1e12,
            "isolated": {"streams": 1, "mean_launch_ms": iso_ms, "achieved": iso,
                         "frac": iso / PEAK_FP32_TFLOPS, "launches_timed": st1.n_pair_launch,
                         "note": "the same launch alone on one stream, 400 MC steps after the timed region"},
        },
    }
    # the same job in the reference's plain shape: dumps of 200 steps, each
    # followed by a blocking copy of its accept records to the host
    mc_ref = g.make_mc(steps, steps_per_dump=200, steps_per_exchange=STEPS_PER_EXCHANGE,
                       move_mode="gaussian", seed=SEED, axes=AXES)
    t0 = time.perf_counter()
    dev.dock_stream(mc_ref, None, blocking=True, stats=False)
    wall_ref = time.perf_counter() - t0
    line["reference_shape"] = {
        "value": job.n_rep * steps / wall_ref, "unit": "replica-steps/s",
        "jobs_per_s": (steps / float(C2_JOB_STEPS)) / wall_ref,
        "shape": "the same job and steps, dumps of 200 steps each followed by a blocking copy of "
                 "the dump's accept records to the host (backend_gpu.cu:96-118)"}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(job, note="; the reference backend's fixed capacities "
                                            "(2048 protein points) cannot hold this complex, so its "
                                            "restatement is timed")
    else:
        line["cpu_baseline"] = None
    dev.close()
    if not args.no_extra:
        print("[bench] headline done; C1 in full", file=sys.stderr, flush=True)
        line["c1"] = c1_extra()
        print("[bench] C1 done; C3 (256 ligands x 5000 steps, batched)", file=sys.stderr, flush=True)
        line["c3"] = batched("c3", 256, 5000, 0, 32, cpu=not args.no_cpu)
        print("[bench] C3 done; 256 C1-shaped complexes batched", file=sys.stderr, flush=True)
        line["c1_batched"] = batched("c1b", 256, 3000, 0, 16, cpu=False)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
