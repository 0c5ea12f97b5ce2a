"""The batched and served configurations at their real shapes, with short step
counts (the full-length runs are bench.py's):

  C3  one 4,000-point target (11 conformers) screened against 256 ligands of
      12-80 atoms (seed 3), 10 ligand conformers x 11 x 4 temperatures = 440
      replicas each (112,640 in all), in one batched launch sequence;
  C4  64 targets with P ~ logU[1,500, 9,000] x 11 conformers, 4 ligands each
      (256 of the configuration's 2,048 complexes), 5 x 11 x 4 = 220 replicas;
  C5  a Poisson stream of 24 requests at 50, 200 and 800 arrivals/s, each one
      protein of 1,500-9,000 points and one ligand of 12-80 atoms with 5-20
      ligand conformers x 11 x 4 replicas.

Each checks, on a sample of jobs, that the job's bits equal its own plain
one-job run (at the tiling shared launches give it), and that the final
poses' energies (gd_energy_eval) meet the FP64 bar of the oracle."""
import ctypes as C
import time

import numpy as np
import pytest

import geauxdock_amd as g
from geauxdock_amd import abi
from tests.oracle_ffi import rel_err

pytestmark = pytest.mark.gpu
AXES = ("temp", "lig", "prt")
TOL = 1e-5


def _batch(devs, mcs):
    outs = [g.host.ReplicaOut(d.n_rep) for d in devs]
    arr = (C.c_void_p * len(devs))(*[d.h.value for d in devs])
    mca = (abi.gd_mc_param * len(devs))(*mcs)
    oa = (abi.gd_replica_out * len(devs))(*[o.c for o in outs])
    g.check(g.lib().gd_batch_dock(arr, len(devs), mca, oa, None), "gd_batch_dock")
    return outs


def _plain(host, mc):
    """The job alone, one job per launch, at its shared-launch tiling."""
    d = g.DeviceJob(host)
    lc, pt, _ = g.host.tiling_lookup(host.desc.lig_natom, host.desc.prt_npoint, host.n_rep, True)
    d.set_tiling(lc, pt)
    out, _ = d.dock(mc, stats=False)
    d.close()
    return out


def _same(a, b):
    for f in ("energy", "movematrix", "idx_tmp", "n_accept", "n_swap_lig", "n_swap_prt", "best_energy"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def _final_pose_parity(host, out, oracle, n=32, seed=0):
    """gd_energy_eval at n final poses of the run against the FP64 oracle at
    the same coordinates; the stored terms equal the fresh evaluation."""
    d = host.desc
    live = np.flatnonzero(out.n_accept > 0)
    if live.size == 0:
        live = np.arange(host.n_rep)
    reps = np.random.default_rng(seed).choice(live, min(n, live.size), replace=False)
    conf = np.stack([reps % d.n_lig, reps // (d.n_tmp * d.n_lig)], 1).astype(np.int32)
    pose = np.ascontiguousarray(out.movematrix[reps], np.float32)
    dev = g.DeviceJob(host)
    lc, pt, _ = g.host.tiling_lookup(d.lig_natom, d.prt_npoint, host.n_rep, True)
    dev.set_tiling(lc, pt)
    terms, raw, coords = dev.eval(conf, pose, want_coords=True)
    dev.close()
    _, r_raw, r_scale, _ = oracle.energy(host, conf, pose, mode=2, coords=coords)
    err = rel_err(raw, r_raw, r_scale).max()
    assert err <= TOL, err
    acc = out.n_accept[reps] > 0
    assert np.array_equal(terms[acc], out.energy[reps][acc])
    return err


def test_c3_screen_at_shape(ene, oracle):
    rng = np.random.default_rng(3)
    ligs = rng.integers(12, 81, 256)
    hosts = [g.synth_screen(4000, 3, int(L), 100 + i, 10, 11, 4, ene=ene) for i, L in enumerate(ligs)]
    mcs = [g.make_mc(20, steps_per_exchange=10, move_mode="gaussian", seed=5, job_id=i, axes=AXES)
           for i in range(len(hosts))]
    devs = [g.DeviceJob(h) for h in hosts]
    assert sum(h.n_rep for h in hosts) == 112640
    outs = _batch(devs, mcs)
    for d in devs:
        d.close()
    for k in (0, 97, 255):
        _same(_plain(hosts[k], mcs[k]), outs[k])
    errs = [_final_pose_parity(hosts[k], outs[k], oracle, seed=k) for k in (5, 200)]
    print("\n[C3] 256 ligands x 440 replicas, 20 steps: final-pose max rel err %.2e" % max(errs))


def test_c4_heterogeneous_at_shape(ene, oracle):
    rng = np.random.default_rng(4)
    hosts = []
    for p in range(64):
        P = int(round(np.exp(rng.uniform(np.log(1500), np.log(9000)))))
        for k in range(4):
            hosts.append(g.synth_screen(P, 1000 + p, int(rng.integers(12, 81)), 5000 + 4 * p + k, 5, 11,
                                        4, ene=ene))
    mcs = [g.make_mc(20, steps_per_exchange=10, move_mode="gaussian", seed=6, job_id=i, axes=AXES)
           for i in range(len(hosts))]
    devs = [g.DeviceJob(h) for h in hosts]
    outs = _batch(devs, mcs)
    for d in devs:
        d.close()
    sizes = [h.desc.prt_npoint for h in hosts]
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes)), 128):
        _same(_plain(hosts[k], mcs[k]), outs[k])
    errs = [_final_pose_parity(hosts[k], outs[k], oracle, seed=k)
            for k in (int(np.argmax(sizes)), 77)]
    print("\n[C4] 256 complexes, P %d-%d: final-pose max rel err %.2e" % (min(sizes), max(sizes), max(errs)))


def test_c5_poisson_serving_at_shape(ene, oracle):
    rng = np.random.default_rng(5)
    hosts, mcs = [], []
    for i in range(24):
        P = int(round(np.exp(rng.uniform(np.log(1500), np.log(9000)))))
        hosts.append(g.synth_complex(P, int(rng.integers(12, 81)), int(rng.integers(5, 21)), 11, 4,
                                     seed=500 + i, ene=ene))
        mcs.append(g.make_mc(int(rng.integers(100, 301)), steps_per_exchange=50, move_mode="gaussian",
                             seed=7, job_id=i, axes=AXES))
    clock = g.lib().gd_server_clock
    clock.restype = C.c_double
    for lam in (50, 200, 800):
        arrive = np.cumsum(np.random.default_rng(lam).exponential(1.0 / lam, len(hosts)))
        cfg = abi.gd_server_config(hbm_budget=200 << 30, max_active=64, sweep_steps=25)
        s = C.c_void_p()
        g.check(g.lib().gd_server_create(C.byref(cfg), C.byref(s)), "gd_server_create")
        outs = [g.host.ReplicaOut(h.n_rep) for h in hosts]
        res = {}
        try:
            t0 = time.perf_counter()
            sub = 0
            tickets = {}
            while len(res) < len(hosts):
                now = time.perf_counter() - t0
                while sub < len(hosts) and arrive[sub] <= now:
                    t = C.c_int64()
                    g.check(g.lib().gd_server_submit(s, C.byref(hosts[sub].desc), C.byref(mcs[sub]),
                                                     C.byref(outs[sub].c), C.byref(t)), "submit")
                    tickets[t.value] = sub
                    sub += 1
                na, nd = C.c_int(), C.c_int()
                g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")
                rr = (abi.gd_job_result * 32)()
                n = C.c_int()
                g.lib().gd_server_results(s, rr, 32, C.byref(n))
                for k in range(n.value):
                    res[tickets[rr[k].ticket]] = (rr[k].status, rr[k].t_submit, rr[k].t_done)
                if na.value == 0 and sub < len(hosts):
                    time.sleep(max(0.0, arrive[sub] - (time.perf_counter() - t0)))
        finally:
            g.lib().gd_server_destroy(s)
        assert all(v[0] == 0 for v in res.values())
        lat = np.array([v[2] - v[1] for v in res.values()])
        print("\n[C5] rate %d/s: %d requests, latency p50 %.3f s p99 %.3f s"
              % (lam, len(res), np.percentile(lat, 50), np.percentile(lat, 99)))
        for k in (0, 11, 23):
            _same(_plain(hosts[k], mcs[k]), outs[k])
        if lam == 200:
            _final_pose_parity(hosts[7], outs[7], oracle, seed=7)
