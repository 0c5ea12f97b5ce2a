"""GPU Monte Carlo path against the CPU oracle: initial pass, trajectories,
acceptance statistics, temperature exchange, conformer pose swaps, records,
batching (persistent work queue + graph-captured segments) and the serving
loop.

Trajectory parity: both sides draw the same Philox numbers, so the accept
decisions agree step for step until a rounding-level energy difference lands
on the Metropolis threshold; after that the chains decorrelate and only
their statistics (acceptance ratio, best-energy distribution) are compared.
"""
import math

import numpy as np
import pytest

import geauxdock_amd as g
from geauxdock_amd import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_job(ene):
    # C1-like shape, reduced protein: 5 ligand x 5 protein conformers x 2 temperatures
    return g.synth_complex(600, 24, 5, 5, 2, floor_temp=0.3, ceiling_temp=0.6, seed=1, ene=ene)


def test_initial_pass_matches_oracle(golden_job, oracle):
    dev = g.DeviceJob(golden_job)
    mc = g.make_mc(0, move_mode="reference")
    out, _ = dev.dock(mc)
    o_out, _, _ = oracle.dock(golden_job, mc, threads=4)
    # every replica starts at its conformer's native pose: forced accept of the initial energy
    assert np.all(out.is_accept == 1)
    err = np.abs(out.energy - o_out.energy).max(0)
    assert err.max() < 5e-5, err
    # replica 0 reproduces the published line
    golden = [0.6434, -0.0367, -0.2079, -0.1845, 0.8521, -0.8880, 0.0523, 0.1744, -1.0, 0.7735]
    assert " ".join("%.4f" % x for x in out.energy[0]) == " ".join("%.4f" % x for x in golden)


@pytest.mark.parametrize("mode", ["reference", "gaussian"])
def test_short_trajectory_matches_oracle(small_job, oracle, mode):
    dev = g.DeviceJob(small_job)
    mc = g.make_mc(30, move_mode=mode, seed=7)
    out, _ = dev.dock(mc)
    o_out, _, _ = oracle.dock(small_job, mc, threads=8)
    same = (out.n_accept == o_out.n_accept) & (np.abs(out.movematrix - o_out.movematrix).max(1) < 1e-3)
    assert same.mean() >= 0.9, same.mean()
    # where the chains agree the accepted energies agree to FP32 rounding
    e = np.abs(out.energy[same] - o_out.energy[same]).max()
    assert e < 1e-3, e


@pytest.mark.parametrize("P,L", [(8, 2), (37, 3)])
def test_small_ragged_trajectory(ene, oracle, P, L):
    """A ligand of 2 or 3 atoms (one padded chunk) on a protein of 8 or 37
    points, with exchange on every axis, follows the oracle's chains."""
    job = g.synth_complex(P, L, 3, 2, 3, floor_temp=0.05, ceiling_temp=0.3, seed=P, ene=ene)
    dev = g.DeviceJob(job)
    mc = g.make_mc(30, steps_per_exchange=5, move_mode="gaussian", seed=P + 1, axes=("temp", "lig", "prt"))
    out, _ = dev.dock(mc)
    o_out, _, _ = oracle.dock(job, mc, threads=8)
    same = (out.n_accept == o_out.n_accept) & (np.abs(out.movematrix - o_out.movematrix).max(1) < 1e-3)
    assert same.mean() >= 0.9, same.mean()
    e = np.abs(out.energy[same] - o_out.energy[same]).max()
    assert e < 1e-3, e


def test_acceptance_statistics_match_oracle(ene, oracle):
    job = g.synth_complex(400, 16, 4, 4, 4, floor_temp=0.05, ceiling_temp=0.5, seed=3, ene=ene)
    dev = g.DeviceJob(job)
    mc = g.make_mc(400, move_mode="gaussian", seed=11)
    out, _ = dev.dock(mc)
    o_out, _, _ = oracle.dock(job, mc, threads=8)
    n = 400 * job.n_rep
    pa, pb = out.n_accept.sum() / n, o_out.n_accept.sum() / n
    se = math.sqrt(pa * (1 - pa) / n + pb * (1 - pb) / n) * 4  # chains are correlated: 4x margin
    assert abs(pa - pb) < 4 * se + 1e-3, (pa, pb)
    from scipy import stats
    ks = stats.ks_2samp(out.best_energy, o_out.best_energy)
    assert ks.pvalue > 1e-3, ks


def test_exchange(ene, oracle):
    job = g.synth_complex(400, 16, 3, 2, 6, floor_temp=0.004, ceiling_temp=0.04, seed=5, ene=ene)
    dev = g.DeviceJob(job)
    mc = g.make_mc(200, steps_per_exchange=10, move_mode="gaussian", seed=13)
    out, _ = dev.dock(mc)
    nl, np_, nt = 3, 2, 6
    # each (ligand conformer, protein conformer) ladder holds every temperature once
    for p in range(np_):
        for l in range(nl):
            reps = [nt * nl * p + nl * t + l for t in range(nt)]
            assert sorted(out.idx_tmp[reps]) == list(range(nt))
    assert out.n_exchange.sum() > 0
    assert out.n_exchange.sum() % 2 == 0
    o_out, _, _ = oracle.dock(job, mc, threads=8)
    # swap counts agree statistically with the oracle's chain
    a, b = out.n_exchange.sum(), o_out.n_exchange.sum()
    assert abs(a - b) <= 6 * math.sqrt(max(a, b)) + 4, (a, b)


def test_swap_identical_conformers(ene):
    """Identical conformers: a pose swap changes no energy, so every attempt is
    accepted and the swap counts equal the pairing rule's attempt counts
    exactly (pairing is bit-exact); temperatures stay put; each replica's
    stored terms are, bit for bit, those gd_energy_eval gives for its final
    pose."""
    from tests.test_oracle import same_conformers, swap_attempts
    nl, np_, nt = 4, 3, 2
    job = g.synth_complex(600, 24, nl, np_, nt, floor_temp=0.05, ceiling_temp=0.2, seed=4, ene=ene)
    same_conformers(job)
    dev = g.DeviceJob(job)
    mc = g.make_mc(40, steps_per_exchange=5, move_mode="gaussian", seed=3, axes=("lig", "prt"))
    out, _ = dev.dock(mc)
    r = np.arange(job.n_rep)
    l, p = r % nl, r // (nl * nt)
    assert np.array_equal(out.n_swap_lig, swap_attempts(nl, 8)[l])
    assert np.array_equal(out.n_swap_prt, swap_attempts(np_, 8)[p])
    assert np.array_equal(out.idx_tmp, (r // nl) % nt) and out.n_exchange.sum() == 0
    terms, _, _ = dev.eval(np.stack([l, p], 1).astype(np.int32), out.movematrix)
    assert np.array_equal(terms, out.energy)


def test_swap_matches_oracle(ene, oracle):
    """All three exchange axes: the GPU chain and the oracle's draw the same
    Philox numbers, so moves, temperature swaps and pose swaps agree replica
    for replica until a rounding-level energy difference meets a threshold."""
    job = g.synth_complex(400, 16, 4, 3, 3, floor_temp=0.05, ceiling_temp=0.3, seed=8, ene=ene)
    mc = g.make_mc(30, steps_per_exchange=5, move_mode="gaussian", seed=17,
                   axes=("temp", "lig", "prt"))
    out, _ = g.DeviceJob(job).dock(mc)
    ref, _, _ = oracle.dock(job, mc, threads=8)
    assert out.n_swap_lig.sum() > 0 and out.n_swap_prt.sum() > 0 and out.n_exchange.sum() > 0
    same = ((out.n_accept == ref.n_accept) & (out.n_swap_lig == ref.n_swap_lig) &
            (out.n_swap_prt == ref.n_swap_prt) & (out.idx_tmp == ref.idx_tmp) &
            (np.abs(out.movematrix - ref.movematrix).max(1) < 1e-3))
    assert same.mean() >= 0.8, same.mean()
    e = np.abs(out.energy[same] - ref.energy[same]).max()
    assert e < 1e-3, e


def test_swap_statistics_match_oracle(ene, oracle):
    job = g.synth_complex(400, 16, 4, 4, 2, floor_temp=0.05, ceiling_temp=0.3, seed=9, ene=ene)
    mc = g.make_mc(300, steps_per_exchange=10, move_mode="gaussian", seed=23, axes=("lig", "prt"))
    out, _ = g.DeviceJob(job).dock(mc)
    ref, _, _ = oracle.dock(job, mc, threads=8)
    for f in ("n_swap_lig", "n_swap_prt"):
        a, b = getattr(out, f).sum(), getattr(ref, f).sum()
        assert a % 2 == 0 and b % 2 == 0
        assert abs(a - b) <= 6 * math.sqrt(max(a, b)) + 4, (f, a, b)


def test_records(small_job):
    dev = g.DeviceJob(small_job)
    mc = g.make_mc(60, steps_per_dump=20, move_mode="reference", seed=3, record=True)
    out, _ = dev.dock(mc)
    rec, cnt = dev.records(20)
    # the last dump covers steps 40..59; a record is the state before a step
    # whose predecessor was accepted
    assert np.all(cnt <= 20) and cnt.sum() > 0
    for r in range(small_job.n_rep):
        steps = rec[r, :cnt[r], 0]
        assert np.all(np.diff(steps) > 0) and np.all((steps >= 40) & (steps < 60))


def _stream(dev, mc, blocking):
    dumps = []
    out, _ = dev.dock_stream(mc, lambda k, s0, rec, cnt: dumps.append((k, s0, rec, cnt)),
                             blocking=blocking)
    return out, dumps


def _flatten(dumps, n_rep):
    """Per replica, the records of every dump in order: list of [n][17]."""
    per = [[] for _ in range(n_rep)]
    for _, _, rec, cnt in dumps:
        for r in range(n_rep):
            per[r].extend(rec[r, :cnt[r]])
    return [np.array(x, np.float32).reshape(-1, 17) for x in per]


def test_record_stream_against_oracle(small_job, oracle):
    """Every dump's records streamed to the host: 90 steps in dumps of 20 (the
    last one 10 steps) with pose swaps and temperature exchange every 15.
    Streamed and blocking deliveries give the same bits; the last dump equals
    gd_dock_records; each record is a true (pose, energy) pair (a fresh
    evaluation gives the same bits); against the oracle's records of the same
    draws, the replicas whose chains agree have the same record steps, poses
    within 1e-5 (the move draws use each side's own libm, a few ulps per
    accepted move) and terms
    within 5e-5 -- the oracle's FP32 chain sums each term serially over every
    pair, and that sum alone is ~2e-5 off FP64 on the PMF term (the records'
    own energies are held to the FP64 bar through the fresh evaluation)."""
    dev = g.DeviceJob(small_job)
    mc = g.make_mc(90, steps_per_dump=20, steps_per_exchange=15, move_mode="gaussian", seed=3,
                   axes=("temp", "lig", "prt"))
    out_a, da = _stream(dev, mc, blocking=False)
    out_b, db = _stream(dev, mc, blocking=True)
    assert [d[:2] for d in da] == [(k, 20 * k) for k in range(5)]
    assert [d[:2] for d in db] == [d[:2] for d in da]
    for x, y in zip(da, db):
        assert np.array_equal(x[3], y[3])
        for r in range(small_job.n_rep):
            assert np.array_equal(x[2][r, :x[3][r]], y[2][r, :y[3][r]])
    assert np.array_equal(out_a.energy, out_b.energy)
    rec, cnt = dev.records(20)
    k, _, rl, cl = da[-1]
    assert np.array_equal(cnt, cl)
    for r in range(small_job.n_rep):
        assert np.array_equal(rec[r, :cnt[r]], rl[r, :cl[r]])
        assert np.all(cl[r] <= 10)
    gpu = _flatten(da, small_job.n_rep)
    # true (pose, energy) pairs, bit for bit
    d = small_job.desc
    conf, pose, want = [], [], []
    for r in range(small_job.n_rep):
        for row in gpu[r]:
            conf.append((r % d.n_lig, r // (d.n_tmp * d.n_lig)))
            pose.append(row[1:7])
            want.append(row[7:17])
    assert len(pose) > 0
    terms, _, _ = dev.eval(np.array(conf, np.int32), np.array(pose, np.float32))
    assert np.array_equal(terms, np.array(want, np.float32))
    # the oracle's records over the whole run
    o_out, _, _, (o_rec, o_cnt) = oracle.dock(small_job, mc, threads=8, records=True)
    same = 0
    for r in range(small_job.n_rep):
        ours = gpu[r]
        theirs = o_rec[r, :o_cnt[r]]
        steps = np.all(np.diff(ours[:, 0]) > 0) if len(ours) > 1 else True
        assert steps and np.all((ours[:, 0] >= 0) & (ours[:, 0] < 90))
        if len(ours) == len(theirs) and np.array_equal(ours[:, 0], theirs[:, 0]):
            dm = np.abs(ours[:, 1:7] - theirs[:, 1:7])
            assert dm.max() <= 1e-5, (r, dm.max())
            rel = np.abs(ours[:, 7:] - theirs[:, 7:]) / np.maximum(np.abs(theirs[:, 7:]), 1.0)
            assert rel.max() <= 5e-5, (r, rel.max())
            same += 1
    assert same >= 0.9 * small_job.n_rep, same


def test_record_csv_sink(small_job, tmp_path):
    """The PrintRecord-format file sink (util_print.C:70-151): one line per
    record of replicas 0..3, '%3d %d' then ' %.4f' per term."""
    dev = g.DeviceJob(small_job)
    mc = g.make_mc(40, steps_per_dump=20, move_mode="gaussian", seed=5)
    out, dumps = _stream(dev, mc, blocking=False)
    path = tmp_path / "rec.txt"
    dev.dock_to_csv(mc, str(path), rep_lo=0, rep_hi=3, header=True)
    lines = path.read_text().splitlines()
    assert lines[0] == "rep step vdw ele pmf psp hdb hpc kde lhm dst total"
    want = []
    for _, _, rec, cnt in dumps:
        for r in range(4):
            for row in rec[r, :cnt[r]]:
                want.append("%3d %d" % (r, int(row[0])) + "".join(" %.4f" % x for x in row[7:17]))
    assert lines[1:] == want


def test_stream_parts_same_bits(ene):
    """Stepping contiguous parts of the replica blocks on 2, 3 and 4 streams
    gives the same bits as one stream: exchange on every axis, records
    streamed, an odd block count (7 blocks: parts of 3+4, 2+2+3, 1+2+2+2).
    One job object switches between stream counts, so the parts' work lists
    are rebuilt in place."""
    job = g.synth_complex(900, 20, 14, 4, 8, seed=23, ene=ene)   # 448 replicas: 7 blocks
    mc = g.make_mc(90, steps_per_dump=40, steps_per_exchange=15, move_mode="gaussian", seed=13,
                   axes=("temp", "lig", "prt"))
    res = []
    dev = g.DeviceJob(job)
    for n in (1, 2, 3, 4, 2):
        dev.set_streams(n)
        dumps = []
        out, st = dev.dock_stream(mc, lambda k, s0, rec, cnt: dumps.append((rec, cnt)), stats=True)
        assert st.n_pair_launch > 0 and st.ms_pair_kernel > 0 and st.ms_step_period > 0
        res.append((out, dumps))
    dev.close()
    a, da = res[0]
    for b, db in res[1:]:
        for f in ("energy", "movematrix", "idx_tmp", "n_accept", "n_exchange", "n_swap_lig", "n_swap_prt",
                  "best_energy"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        assert len(da) == len(db)
        for (ra, ca), (rb, cb) in zip(da, db):
            assert np.array_equal(ca, cb)
            for r in range(job.n_rep):
                assert np.array_equal(ra[r, :ca[r]], rb[r, :cb[r]])


def test_stream_count_bounds(small_job):
    """Stream counts outside 1-4 are refused; more streams than replica
    blocks step the blocks there are (here one), with the same bits."""
    dev = g.DeviceJob(small_job)
    for bad in (0, 5):
        with pytest.raises(g.GDError):
            dev.set_streams(bad)
    mc = g.make_mc(30, steps_per_exchange=10, move_mode="gaussian", seed=4)
    outs = []
    for n in (1, 4):
        dev.set_streams(n)
        outs.append(dev.dock(mc)[0])
    dev.close()
    assert -(-small_job.n_rep // 64) < 4
    for f in ("energy", "movematrix", "n_accept"):
        assert np.array_equal(getattr(outs[0], f), getattr(outs[1], f)), f


def test_dock_is_deterministic(small_job):
    dev = g.DeviceJob(small_job)
    mc = g.make_mc(20, move_mode="gaussian", seed=21)
    a, _ = dev.dock(mc)
    b, _ = dev.dock(mc)
    assert np.array_equal(a.energy, b.energy) and np.array_equal(a.movematrix, b.movematrix)
    # the replica state stays on the device and reads back unchanged
    c = dev.replicas()
    for f in ("movematrix", "energy", "idx_tmp", "n_accept", "n_exchange", "best_energy", "is_accept"):
        assert np.array_equal(getattr(b, f), getattr(c, f)), f


def _batch(jobs, mcs, stats=False):
    import ctypes as C
    outs = [g.host.ReplicaOut(j.n_rep) for j in jobs]
    arr = (C.c_void_p * len(jobs))(*[j.h.value for j in jobs])
    mca = (abi.gd_mc_param * len(jobs))(*mcs)
    oa = (abi.gd_replica_out * len(jobs))(*[o.c for o in outs])
    st = abi.gd_stats()
    g.check(g.lib().gd_batch_dock(arr, len(jobs), mca, oa, C.byref(st) if stats else None),
            "gd_batch_dock")
    return outs


def test_batch_matches_plain_bitwise(ene):
    """A job computes the same bits alone (gd_dock) and inside a batch
    (persistent work-queue pair kernel, graph-replayed segments)."""
    hjobs = [g.synth_complex(P, L, 2, 3, 2, seed=s, ene=ene)
             for P, L, s in [(500, 12, 1), (900, 30, 2), (1600, 20, 3)]]
    axes = [("temp", "lig", "prt"), ("temp",), ("lig",)]
    mcs = [g.make_mc(120, steps_per_exchange=20, move_mode="gaussian", seed=5, job_id=k, axes=axes[k])
           for k in range(3)]
    plain = []
    for h, mc in zip(hjobs, mcs):
        d = g.DeviceJob(h)   # the tiling shared launches give this complex
        lc, pt, _ = g.host.tiling_lookup(h.desc.lig_natom, h.desc.prt_npoint, h.n_rep, True)
        d.set_tiling(lc, pt)
        plain.append(d.dock(mc, stats=False)[0])
    devs = [g.DeviceJob(h) for h in hjobs]
    for d in devs:   # a tiling the batch must replace by the table's
        d.set_tiling(8, d.tiling()[1])
    for rep in range(3):  # several graph replays, eager work in between
        outs = _batch(devs, mcs)
        for a, b in zip(plain, outs):
            assert np.array_equal(a.energy, b.energy)
            assert np.array_equal(a.movematrix, b.movematrix)
            assert np.array_equal(a.idx_tmp, b.idx_tmp)
            assert np.array_equal(a.n_swap_lig, b.n_swap_lig)
            assert np.array_equal(a.n_swap_prt, b.n_swap_prt)
        devs[0].eval(np.zeros((1, 2), np.int32), np.zeros((1, 6), np.float32))
    for h, d in zip(hjobs, devs):  # the batch ran every job at its table tiling
        lc, pt, _ = g.host.tiling_lookup(h.desc.lig_natom, h.desc.prt_npoint, h.n_rep, True)
        assert d.tiling() == (lc, pt)


def test_batch_timed_path_matches_graph_path(ene):
    hjobs = [g.synth_complex(700, 16, 2, 2, 1, seed=s, ene=ene) for s in (4, 5)]
    mcs = [g.make_mc(100, steps_per_exchange=20, move_mode="gaussian", seed=9, job_id=k,
                     axes=("lig", "prt")) for k in range(2)]
    devs = [g.DeviceJob(h) for h in hjobs]
    a = _batch(devs, mcs, stats=False)
    b = _batch(devs, mcs, stats=True)
    for x, y in zip(a, b):
        assert np.array_equal(x.energy, y.energy)


def test_server_alone_vs_mix(ene):
    """Continuous batching: a request admitted while others run returns the
    same result as when it runs alone."""
    import ctypes as C
    target = g.synth_complex(800, 20, 3, 2, 2, seed=31, ene=ene)
    others = [g.synth_complex(P, L, 2, 2, 2, seed=s, ene=ene)
              for P, L, s in [(1500, 40, 32), (600, 12, 33), (1200, 60, 34)]]
    mc_t = g.make_mc(150, steps_per_exchange=25, move_mode="gaussian", seed=77, job_id=1000,
                     axes=("temp", "lig", "prt"))

    def serve(jobs_and_mc, sweep):
        cfg = abi.gd_server_config(hbm_budget=64 << 30, max_active=8, sweep_steps=sweep)
        s = C.c_void_p()
        g.check(g.lib().gd_server_create(C.byref(cfg), C.byref(s)), "gd_server_create")
        outs = {}
        try:
            pending = list(jobs_and_mc)
            done = 0
            while done < len(jobs_and_mc):
                if pending:  # one arrival per sweep: later jobs join a running batch
                    h, mc, key = pending.pop(0)
                    o = g.host.ReplicaOut(h.n_rep)
                    t = C.c_int64()
                    g.check(g.lib().gd_server_submit(s, C.byref(h.desc), C.byref(mc), C.byref(o.c),
                                                     C.byref(t)), "submit")
                    outs[key] = o
                na, nd = C.c_int(), C.c_int()
                g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")
                res = (abi.gd_job_result * 16)()
                n = C.c_int()
                g.lib().gd_server_results(s, res, 16, C.byref(n))
                for i in range(n.value):
                    assert res[i].status == 0
                    assert res[i].t_done >= res[i].t_admit >= res[i].t_submit
                done += n.value
        finally:
            g.lib().gd_server_destroy(s)
        return outs

    alone = serve([(target, mc_t, "t")], sweep=40)["t"]
    mix = [(others[0], g.make_mc(300, move_mode="gaussian", seed=1, job_id=1), "a"),
           (others[1], g.make_mc(90, move_mode="gaussian", seed=2, job_id=2), "b"),
           (target, mc_t, "t"),
           (others[2], g.make_mc(200, steps_per_exchange=10, move_mode="gaussian", seed=3, job_id=3), "c")]
    mixed = serve(mix, sweep=17)["t"]
    assert np.array_equal(alone.energy, mixed.energy)
    assert np.array_equal(alone.movematrix, mixed.movematrix)
    assert np.array_equal(alone.idx_tmp, mixed.idx_tmp)
    assert np.array_equal(alone.n_swap_lig, mixed.n_swap_lig)
    assert alone.n_swap_lig.sum() > 0 and alone.n_swap_prt.sum() > 0


def test_server_file_requests(ene, golden_job):
    """Host input pipeline: a request given by its files is parsed by the
    server's loader thread and docks to the same bits as the same complex
    submitted as a parsed descriptor; a request whose files are missing comes
    back as a result with the loader's status and leaves nothing to take."""
    import ctypes as C
    import os
    import time
    from tests.conftest import GOLDEN_DATA as d
    mc = g.make_mc(60, steps_per_exchange=20, move_mode="gaussian", seed=3, job_id=5,
                   axes=("temp", "lig", "prt"))
    cfg = abi.gd_server_config(hbm_budget=16 << 30, max_active=4, sweep_steps=20)
    s = C.c_void_p()
    g.check(g.lib().gd_server_create(C.byref(cfg), C.byref(s)), "gd_server_create")
    try:
        def files(sdf):
            return abi.gd_request_files(sdf.encode(), os.path.join(d, "1a07C1-0.8.ff").encode(),
                                        os.path.join(d, "1a07C-1.pdb").encode(), b"1a07C1",
                                        C.pointer(ene), 1, 0.044, 0.036)
        req, bad = files(os.path.join(d, "1a07C1.sdf")), files(os.path.join(d, "missing.sdf"))
        tf, tb, td = C.c_int64(), C.c_int64(), C.c_int64()
        g.check(g.lib().gd_server_submit_files(s, C.byref(req), C.byref(mc), C.byref(tf)), "submit_files")
        g.check(g.lib().gd_server_submit_files(s, C.byref(bad), C.byref(mc), C.byref(tb)), "submit_files")
        o_desc = g.host.ReplicaOut(golden_job.n_rep)
        g.check(g.lib().gd_server_submit(s, C.byref(golden_job.desc), C.byref(mc), C.byref(o_desc.c),
                                         C.byref(td)), "submit")
        status = {}
        for _ in range(400):
            na, nd = C.c_int(), C.c_int()
            g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")
            res = (abi.gd_job_result * 8)()
            n = C.c_int()
            g.lib().gd_server_results(s, res, 8, C.byref(n))
            for i in range(n.value):
                status[res[i].ticket] = res[i].status
            if len(status) == 3:
                break
            if na.value == 0:
                time.sleep(0.01)  # the loader thread is still parsing
        assert status.get(tf.value) == 0 and status.get(td.value) == 0, status
        assert status.get(tb.value, 0) != 0, status
        n_rep = C.c_int()
        assert g.lib().gd_server_take(s, tf, C.byref(n_rep), None) == abi.GD_OK
        assert n_rep.value == golden_job.n_rep
        o_file = g.host.ReplicaOut(n_rep.value)
        assert g.lib().gd_server_take(s, tf, None, C.byref(o_file.c)) == abi.GD_OK
        for f in ("energy", "movematrix", "idx_tmp", "n_accept", "n_swap_lig", "best_energy"):
            assert np.array_equal(getattr(o_file, f), getattr(o_desc, f)), f
        assert g.lib().gd_server_take(s, tf, C.byref(n_rep), None) == abi.GD_ERR_ARG  # released
        assert g.lib().gd_server_take(s, tb, C.byref(n_rep), None) == abi.GD_ERR_ARG
    finally:
        g.lib().gd_server_destroy(s)


def _shared_bytes(h):
    """Device bytes of a complex at its shared-launch tiling."""
    d = g.DeviceJob(h)
    lc, pt, _ = g.host.tiling_lookup(h.desc.lig_natom, h.desc.prt_npoint, h.n_rep, True)
    d.set_tiling(lc, pt)
    b = g.lib().gd_job_device_bytes(d.h)
    d.close()
    return b


def test_server_admission_budget(ene):
    """A request larger than the whole budget is refused; requests that fit
    only one at a time run one after another."""
    import ctypes as C
    h = g.synth_complex(800, 20, 2, 2, 2, seed=41, ene=ene)
    foot = _shared_bytes(h)
    assert foot > 0
    cfg = abi.gd_server_config(hbm_budget=foot + foot // 2, max_active=8, sweep_steps=10)
    s = C.c_void_p()
    g.check(g.lib().gd_server_create(C.byref(cfg), C.byref(s)), "create")
    try:
        big = g.synth_complex(4000, 40, 8, 4, 4, seed=42, ene=ene)
        t = C.c_int64()
        mc = g.make_mc(20, seed=1)
        assert g.lib().gd_server_submit(s, C.byref(big.desc), C.byref(mc), None, C.byref(t)) == abi.GD_ERR_BUSY
        outs = [g.host.ReplicaOut(h.n_rep) for _ in range(3)]
        for k in range(3):
            g.check(g.lib().gd_server_submit(s, C.byref(h.desc), C.byref(mc), C.byref(outs[k].c),
                                             C.byref(t)), "submit")
        na, nd = C.c_int(), C.c_int()
        g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")
        used, pend = C.c_size_t(), C.c_int()
        g.lib().gd_server_usage(s, C.byref(used), C.byref(pend))
        assert na.value <= 1 and used.value <= cfg.hbm_budget
        total = nd.value
        for _ in range(20):
            g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")
            total = nd.value
            if total == 3:
                break
        assert total == 3
        # identical requests, identical results
        assert np.array_equal(outs[0].energy, outs[1].energy)
    finally:
        g.lib().gd_server_destroy(s)


def test_autotune_keeps_parity(ene, oracle):
    job = g.synth_complex(1200, 24, 2, 2, 2, seed=51, ene=ene)
    dev = g.DeviceJob(job)
    import ctypes as C
    t = abi.gd_tiling()
    ms = C.c_double()
    g.check(g.lib().gd_autotune(dev.h, 3, C.byref(t), C.byref(ms)), "gd_autotune")
    assert t.atoms_per_lane in (4, 8, 16) and t.tile_points > 0 and ms.value > 0
    from tests.oracle_ffi import random_poses, rel_err
    conf, pose = random_poses(job, 32, seed=2)
    _, raw, coords = dev.eval(conf, pose, want_coords=True)
    _, r_raw, r_scale, _ = oracle.energy(job, conf, pose, mode=2, coords=coords)
    assert rel_err(raw, r_raw, r_scale).max() <= 1e-5


def _server(**kw):
    import ctypes as C
    cfg = abi.gd_server_config(**kw)
    s = C.c_void_p()
    g.check(g.lib().gd_server_create(C.byref(cfg), C.byref(s)), "gd_server_create")
    return s


def _submit(s, h, mc, out=None):
    import ctypes as C
    t = C.c_int64()
    g.check(g.lib().gd_server_submit(s, C.byref(h.desc), C.byref(mc), C.byref(out.c) if out else None,
                                     C.byref(t)), "submit")
    return t.value


def _drain(s, results, want, max_steps=200):
    import ctypes as C
    for _ in range(max_steps):
        na, nd = C.c_int(), C.c_int()
        g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")
        res = (abi.gd_job_result * 32)()
        n = C.c_int()
        g.lib().gd_server_results(s, res, 32, C.byref(n))
        for i in range(n.value):
            results[res[i].ticket] = (res[i].status, res[i].replica_steps, res[i].t_submit,
                                      res[i].t_admit, res[i].t_done)
        if len(results) >= want:
            return
    raise AssertionError("server did not finish: %d of %d" % (len(results), want))


def test_server_charges_device_bytes(ene):
    """While jobs are active the server's usage is the sum of their actual
    device bytes (at the shared-launch tiling); when all retire it is 0."""
    import ctypes as C
    hs = [g.synth_complex(P, L, 2, 2, 2, seed=s, ene=ene) for P, L, s in [(700, 14, 61), (1500, 33, 62),
                                                                         (2600, 70, 63)]]
    want = sum(_shared_bytes(h) for h in hs)
    s = _server(hbm_budget=64 << 30, max_active=8, sweep_steps=10)
    try:
        for k, h in enumerate(hs):
            _submit(s, h, g.make_mc(40, seed=1, job_id=k))
        na, nd = C.c_int(), C.c_int()
        g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")
        used, pend = C.c_size_t(), C.c_int()
        g.lib().gd_server_usage(s, C.byref(used), C.byref(pend))
        assert na.value == 3 and pend.value == 0 and used.value == want, (used.value, want)
        res = {}
        _drain(s, res, 3)
        g.lib().gd_server_usage(s, C.byref(used), C.byref(pend))
        assert used.value == 0 and all(v[0] == 0 for v in res.values())
    finally:
        g.lib().gd_server_destroy(s)


def test_server_cancel(ene):
    """Cancelling a pending request, an active job and a file request (queued
    or already with the loader thread) each yields one GD_ERR_STATE result and
    nothing runs on; an unknown ticket is an argument error."""
    import ctypes as C
    import os
    from tests.conftest import GOLDEN_DATA as d
    h = g.synth_complex(900, 20, 2, 2, 2, seed=71, ene=ene)
    s = _server(hbm_budget=64 << 30, max_active=1, sweep_steps=10)
    try:
        ta = _submit(s, h, g.make_mc(200, seed=1, job_id=1))
        tb = _submit(s, h, g.make_mc(200, seed=1, job_id=2))
        na, nd = C.c_int(), C.c_int()
        g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")   # a active, b pending
        req = abi.gd_request_files(os.path.join(d, "1a07C1.sdf").encode(),
                                   os.path.join(d, "1a07C1-0.8.ff").encode(),
                                   os.path.join(d, "1a07C-1.pdb").encode(), b"1a07C1",
                                   C.pointer(ene), 1, 0.044, 0.036)
        tc = C.c_int64()
        g.check(g.lib().gd_server_submit_files(s, C.byref(req), C.byref(g.make_mc(50, seed=2)),
                                               C.byref(tc)), "submit_files")
        assert g.lib().gd_server_cancel(s, tc.value) == abi.GD_OK
        assert g.lib().gd_server_cancel(s, tb) == abi.GD_OK
        assert g.lib().gd_server_cancel(s, ta) == abi.GD_OK
        assert g.lib().gd_server_cancel(s, 987654) == abi.GD_ERR_ARG
        res = {}
        import time
        for _ in range(100):
            _drain(s, res, 0, max_steps=1)
            if len(res) == 3:
                break
            time.sleep(0.02)   # a file request cancelled while being parsed reports when parsed
        assert set(res) == {ta, tb, tc.value}, res
        assert all(v[0] == abi.GD_ERR_STATE for v in res.values()), res
        assert res[ta][1] > 0 and res[tb][1] == 0
        n_rep = C.c_int()
        assert g.lib().gd_server_take(s, tc, C.byref(n_rep), None) == abi.GD_ERR_ARG
        used, pend = C.c_size_t(), C.c_int()
        g.lib().gd_server_usage(s, C.byref(used), C.byref(pend))
        assert used.value == 0 and pend.value == 0
    finally:
        g.lib().gd_server_destroy(s)


def test_server_retires_converged_jobs(ene):
    """With convergence retirement a job leaves after its third sweep when the
    tolerance cannot be missed (3 x 10 steps instead of 200), and its result is
    the same alone and among other jobs."""
    h = g.synth_complex(800, 20, 2, 2, 2, seed=81, ene=ene)
    mc = g.make_mc(200, steps_per_exchange=20, move_mode="gaussian", seed=5, job_id=9,
                   axes=("temp", "lig", "prt"))
    outs = []
    for others in (0, 2):
        s = _server(hbm_budget=64 << 30, max_active=8, sweep_steps=10, converge_sweeps=2,
                    converge_tol=1e9)
        try:
            for k in range(others):
                _submit(s, g.synth_complex(1200, 30, 2, 2, 2, seed=82 + k, ene=ene),
                        g.make_mc(500, seed=3, job_id=k))
            o = g.host.ReplicaOut(h.n_rep)
            t = _submit(s, h, mc, o)
            res = {}
            _drain(s, res, others + 1)
            assert res[t][0] == 0 and res[t][1] == h.n_rep * 30, res[t]
            outs.append(o)
        finally:
            g.lib().gd_server_destroy(s)
    assert np.array_equal(outs[0].energy, outs[1].energy)
    assert np.array_equal(outs[0].movematrix, outs[1].movematrix)


def test_server_submit_does_not_wait_for_a_sweep(ene):
    """A request submitted while gd_server_step runs a long sweep returns at
    once and is stamped on arrival, before that sweep ends."""
    import ctypes as C
    import threading
    import time
    big = g.synth_complex(3000, 40, 20, 11, 8, seed=2, ene=ene)
    small = g.synth_complex(600, 12, 2, 2, 1, seed=91, ene=ene)
    s = _server(hbm_budget=64 << 30, max_active=8, sweep_steps=3000)
    try:
        _submit(s, big, g.make_mc(3000, seed=1, job_id=1))
        clock = g.lib().gd_server_clock
        clock.restype = C.c_double
        span = {}

        def run():
            na, nd = C.c_int(), C.c_int()
            span["t0"] = clock(s)
            g.check(g.lib().gd_server_step(s, C.byref(na), C.byref(nd)), "step")
            span["t1"] = clock(s)

        th = threading.Thread(target=run)
        th.start()
        time.sleep(0.05)
        w0 = time.perf_counter()
        t = _submit(s, small, g.make_mc(10, seed=2, job_id=2))
        dt = time.perf_counter() - w0
        th.join()
        res = {}
        _drain(s, res, 2)
        assert span["t1"] - span["t0"] > 0.2, span   # the sweep was long
        # a submit that waited for the sweep would take the rest of it (over
        # 0.15 s, the submit starting 0.05 s in); one that does not, milliseconds
        assert dt < 0.1, dt
        assert res[t][2] < span["t1"], (res[t], span)
    finally:
        g.lib().gd_server_destroy(s)


def test_tiling_is_a_function_of_the_job(tmp_path):
    """A job's tiling, and so its bits, is the same in a fresh process and in
    one that first autotuned another job of the same size bucket."""
    import subprocess
    import sys
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = r"""
import sys, ctypes as C, numpy as np
sys.path.insert(0, %r)
import geauxdock_amd as g
from geauxdock_amd import abi
g.check(g.lib().gd_init(0), "init")
ene = g.load_enepara()
if sys.argv[1] == "tune":
    other = g.DeviceJob(g.synth_complex(2200, 33, 4, 4, 2, seed=7, ene=ene))
    t = abi.gd_tiling(); ms = C.c_double()
    g.check(g.lib().gd_autotune(other.h, 3, C.byref(t), C.byref(ms)), "tune")
dev = g.DeviceJob(g.synth_complex(2100, 35, 4, 4, 2, seed=8, ene=ene))
out, _ = dev.dock(g.make_mc(40, seed=3))
np.save(sys.argv[2], np.concatenate([out.energy.ravel(), out.movematrix.ravel(),
                                     np.array(dev.tiling(), np.float32)]))
""" % root
    res = []
    for mode in ("plain", "tune"):
        f = str(tmp_path / (mode + ".npy"))
        r = subprocess.run([sys.executable, "-c", code, mode, f], capture_output=True, text=True,
                           timeout=240)
        assert r.returncode == 0, r.stderr
        res.append(np.load(f))
    assert np.array_equal(res[0], res[1])


def test_autotune_shared(ene):
    """The shared-launch sweep returns a candidate of its grid and leaves the
    jobs at the table's tiling."""
    import ctypes as C
    hs = [g.synth_complex(P, L, 2, 3, 2, seed=s, ene=ene) for P, L, s in [(1800, 20, 95), (2500, 44, 96)]]
    devs = [g.DeviceJob(h) for h in hs]
    arr = (C.c_void_p * len(devs))(*[d.h.value for d in devs])
    t = abi.gd_tiling()
    ms = C.c_double()
    g.check(g.lib().gd_autotune_shared(arr, len(devs), 4, C.byref(t), C.byref(ms)), "tune shared")
    assert t.atoms_per_lane == 4 and t.tile_points in (32, 40, 48, 64, 80, 96, 128)
    assert t.waves_per_cu in (16, 24, 32) and ms.value > 0
    for h, d in zip(hs, devs):
        lc, pt, _ = g.host.tiling_lookup(h.desc.lig_natom, h.desc.prt_npoint, h.n_rep, True)
        assert d.tiling() == (lc, pt)
