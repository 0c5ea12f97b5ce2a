"""Pins of the CPU oracle (CPU only).

Each check is chosen so a plausible mistake fails it: the reference's golden
line and fixture sizes, the Random123 known-answer vectors, a bit-for-bit
comparison with the reference CPU backend compiled from its own sources
(where oracle/_ref was built), closed forms of each pair term, rigid-motion
invariance, and a brute-force numpy evaluation of a tiny complex.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import geauxdock_amd as g
from tests.oracle_ffi import REF, Oracle, random_poses, rel_err

GOLDEN = [0.6434, -0.0367, -0.2079, -0.1845, 0.8521, -0.8880, 0.0523, 0.1744, -1.0000, 0.7735]

# src/module/Random123/examples/kat_vectors:27-29 (philox4x32, 10 rounds) and :24-26 (7 rounds)
KAT = [
    (10, [0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    (10, [0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    (10, [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    (7, [0, 0, 0, 0], [0, 0], [0x5f6fb709, 0x0d893f64, 0x4f121f81, 0x4f730a48]),
    (7, [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0x4dfccaba, 0x190a87f0, 0xc47362ba, 0xb6b5242a]),
]


@pytest.mark.parametrize("rounds,ctr,key,expect", KAT)
def test_philox_kat_oracle(oracle, rounds, ctr, key, expect):
    assert oracle.philox(ctr, key, rounds) == expect


@pytest.mark.parametrize("rounds,ctr,key,expect", KAT)
def test_philox_kat_library(rounds, ctr, key, expect):
    """The library's host/device Philox header, called on the host side."""
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    g.lib().gd_philox(c, k, rounds, o)
    assert list(o) == expect


def test_step_draw_layout(oracle):
    # slot 0 -> draws 0..3, slot 1 -> draws 4..7 of counter (step, replica, job, slot)
    seed, job, rep, step = 0x1234567890ABCDEF, 7, 99, 12345
    w = oracle.draws(seed, job, rep, step)
    key = [seed & 0xffffffff, seed >> 32]
    assert w[:4] == oracle.philox([step, rep, job, 0], key)
    assert w[4:] == oracle.philox([step, rep, job, 1], key)


def test_golden_line_all_modes(golden_job, oracle):
    conf = np.zeros((1, 2), np.int32)
    pose = np.zeros((1, 6), np.float32)
    for mode in (0, 1, 2):
        t, _, _, _ = oracle.energy(golden_job, conf, pose, mode=mode)
        assert " ".join("%.4f" % x for x in t[0]) == " ".join("%.4f" % x for x in GOLDEN)


def test_fixture_sizes(golden_job):
    # doc/instructions.txt:189: n_lig=51, lig_natom=34, prt_npoint=344, kde_npoint=454, mcs_nrow=11
    d = golden_job.desc
    assert (d.n_lig, d.n_prt, d.n_tmp) == (51, 1, 1)
    assert (d.lig_natom, d.prt_npoint, d.kde_npoint, d.mcs_nrow) == (34, 344, 454, 11)


def test_fixture_kde_bins_and_mcs(golden_job):
    a = golden_job.arrays()
    names = {2: "C.2", 3: "C.3", 4: "C.ar", 13: "N.am", 16: "O.2", 17: "O.3", 18: "O.co2", 19: "P.3"}
    t, n = np.unique(a["kde_t"], return_counts=True)
    got = {names[int(x)]: int(c) for x, c in zip(t, n)}
    # counted from the KDE lines of 1a07C1-0.8.ff
    assert got == {"C.2": 74, "C.3": 129, "C.ar": 114, "N.am": 39, "O.2": 42, "O.3": 9,
                   "O.co2": 44, "P.3": 3}
    assert np.all(np.diff(a["kde_t"]) >= 0)
    ncol = (a["mcs_x"] != np.float32(g.abi.GD_MCS_INVALID)).sum(1)
    assert list(ncol) == [15, 15, 6, 20, 15, 13, 15, 6, 15, 15, 13]


def test_fixture_conformer_rmsd(golden_job):
    """The loader's superposition RMSD reproduces the SDF's own ENSEMBLE_RMSD list."""
    raw = C.c_int()
    rm = (C.c_float * 64)()
    g.lib().gd_host_job_diag(golden_job._h, C.byref(raw), rm, 64, None, 0)
    assert raw.value == 50
    with open(os.path.join(os.path.dirname(__file__), "golden", "data", "1a07C1.sdf")) as f:
        lines = f.read().splitlines()
    i = [k for k, s in enumerate(lines) if "ENSEMBLE_RMSD" in s][0]
    ref = np.array([float(x) for x in lines[i + 1].split()])
    assert np.abs(np.array(rm[:50]) - ref).max() < 1.5e-3


def test_kabsch_invariants():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(20, 3)).astype(np.float32)
    a -= a.mean(0)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    b = (a @ R.T).astype(np.float32)
    f = lambda x: x.ctypes.data_as(C.POINTER(C.c_float))
    assert g.lib().gd_kabsch_rmsd(20, f(a), f(b)) < 1e-4          # pure rotation
    c = (a + 0.1 * rng.normal(size=a.shape)).astype(np.float32)
    c -= c.mean(0)
    rms = math.sqrt(((a - c) ** 2).sum() / 20)
    assert g.lib().gd_kabsch_rmsd(20, f(a), f(c)) <= rms + 1e-5     # superposition can only help


def test_temperature_ladder():
    mb = (C.c_float * 8)()
    g.lib().gd_set_temperature(8, C.c_float(0.004), C.c_float(0.04), mb)
    T = -1.0 / np.array(mb[:])
    assert abs(T[0] - 0.004) < 1e-9 and abs(T[-1] - 0.04) < 1e-6
    assert np.allclose(T[1:] / T[:-1], 10 ** (1 / 7), rtol=1e-5)
    mb1 = (C.c_float * 1)()
    g.lib().gd_set_temperature(1, C.c_float(0.044), C.c_float(0.036), mb1)
    assert abs(mb1[0] + 1 / 0.044) < 1e-3


def test_exp_underflow_threshold(oracle):
    """The forced initial accept fails below this float (expf underflows)."""
    x = oracle.lib.or_exp_underflow_threshold()
    assert x == np.float32(-103.97207641601562)
    assert math.exp(x) > 0 and np.float32(math.exp(np.nextafter(np.float32(x), np.float32(-200)))) == 0


@pytest.mark.skipif(not os.path.exists(REF), reason="reference backend not built (oracle/ref/Makefile)")
def test_literal_mode_bit_exact_vs_reference_backend(golden_job, oracle):
    """FP32 'literal' mode == the reference CPU backend compiled from its sources."""
    ref = C.CDLL(REF)
    conf, pose = random_poses(golden_job, 200, seed=1, trans=1.0, rot=3.0)
    t, _, _, _ = oracle.energy(golden_job, conf, pose, mode=1)
    rt = np.zeros((200, 10), np.float32)
    fd = os.dup(1)
    os.dup2(os.open(os.devnull, os.O_WRONLY), 1)  # the reference Dock prints its report
    try:
        rc = ref.ref_energy(C.byref(golden_job.desc), 200, conf.ctypes.data_as(C.POINTER(C.c_int)),
                            pose.ctypes.data_as(C.POINTER(C.c_float)),
                            rt.ctypes.data_as(C.POINTER(C.c_float)))
    finally:
        os.dup2(fd, 1)
    assert rc == 0
    assert np.array_equal(t.astype(np.float32), rt)


def test_fp32_modes_vs_fp64(golden_job, oracle):
    conf, pose = random_poses(golden_job, 64, seed=2)
    _, r64, s64, _ = oracle.energy(golden_job, conf, pose, mode=2)
    for mode in (0, 1):
        _, r32, _, _ = oracle.energy(golden_job, conf, pose, mode=mode)
        assert rel_err(r32, r64, s64).max() < 2e-5


def test_rigid_motion_invariance(ene, oracle):
    """Every pair term is invariant under one global rigid motion of the system."""
    job = g.synth_complex(300, 10, 2, 2, 1, seed=3, ene=ene)
    a = job.arrays()
    rng = np.random.default_rng(1)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.linalg.det(q))
    t = np.array([3.0, -2.0, 5.0])
    mv = lambda x, y, z: np.einsum("ij,...j->...i", q, np.stack([x, y, z], -1)) + t
    b = dict(a)
    P = mv(a["prt_x"], a["prt_y"], a["prt_z"])
    b["prt_x"], b["prt_y"], b["prt_z"] = P[..., 0], P[..., 1], P[..., 2]
    K = mv(a["kde_x"], a["kde_y"], a["kde_z"])
    b["kde_x"], b["kde_y"], b["kde_z"] = K[..., 0], K[..., 1], K[..., 2]
    Mv = mv(a["mcs_x"], a["mcs_y"], a["mcs_z"])
    inv = a["mcs_x"] == np.float32(g.abi.GD_MCS_INVALID)
    b["mcs_x"] = np.where(inv, a["mcs_x"], Mv[..., 0])
    b["mcs_y"] = np.where(inv, a["mcs_y"], Mv[..., 1])
    b["mcs_z"] = np.where(inv, a["mcs_z"], Mv[..., 2])
    b["lig_center"] = mv(*a["lig_center"])
    b["pocket_center"] = mv(*a["pocket_center"])
    # ligand: express the rotated native frame through the conformer coordinates
    L0 = np.einsum("ij,...j->...i", q, np.stack([a["lig_x"], a["lig_y"], a["lig_z"]], -1))
    b["lig_x"], b["lig_y"], b["lig_z"] = L0[..., 0], L0[..., 1], L0[..., 2]
    job2 = g.HostJob.from_arrays(ene, **b)
    conf = np.array([[0, 0], [1, 1]], np.int32)
    pose = np.zeros((2, 6), np.float32)
    _, r1, s1, _ = oracle.energy(job, conf, pose, mode=2)
    _, r2, s2, _ = oracle.energy(job2, conf, pose, mode=2)
    assert rel_err(r2, r1, s1).max() < 1e-9 + 1e-11 / max(1e-30, s1.min())


def _brute(job, conf, pose):
    """Independent numpy FP64 evaluation straight from the term definitions
    (energy_prt_v1.C, energy_kde_v1.C, energy_mcs_v1.C); small inputs only."""
    a = job.arrays()
    e = job.ene
    out = []
    for (li, pi), m in zip(conf, pose.astype(np.float64)):
        s1, c1 = math.sin(m[3]), math.cos(m[3])
        s2, c2 = math.sin(m[4]), math.cos(m[4])
        s3, c3 = math.sin(m[5]), math.cos(m[5])
        R = np.array([[c1 * c2, c1 * s2 * s3 - c3 * s1, s1 * s3 + c1 * c3 * s2],
                      [c2 * s1, c1 * c3 + s1 * s2 * s3, c3 * s1 * s2 - c1 * s3],
                      [-s2, c2 * s3, c2 * c3]])
        X0 = np.stack([a["lig_x"][li], a["lig_y"][li], a["lig_z"][li]], 1).astype(np.float64)
        X = X0 @ R.T + m[:3] + a["lig_center"]
        Pp = np.stack([a["prt_x"][pi], a["prt_y"][pi], a["prt_z"][pi]], 1).astype(np.float64)
        vdw = ele = pmf = psp = hdb = hpc = 0.0
        for l in range(X.shape[0]):
            lt = a["lig_t"][l]
            h1 = 0.0
            for p in range(Pp.shape[0]):
                pt = a["prt_t"][p]
                r = np.linalg.norm(X[l] - Pp[p])
                if a["prt_cdc"][p] == 1 and r * r <= 81:
                    x = r * r / 81.0
                    h1 += a["prt_hpp"][p] * (1 - (3.5 * x - 4.5 * x ** 2 + 2.5 * x ** 3 - 0.5 * x ** 4))
                p1 = e.p1a[lt][pt] / r ** 9
                p2 = e.p2a[lt][pt] / r ** 6
                vdw += (p1 - p2) / (p1 * e.lj0 * (1 + e.lj1 * r * r) + 1)
                s = e.el1 * r
                gg = e.el0 + e.a1 * s * s + e.b1 * s ** 3 if s < 1 else 1 / s
                ele += a["lig_c"][l] * a["prt_ele"][p] * gg
                pmf += e.pmf1[lt][pt] / (1 + math.exp((-0.5 * r + 6) * (r - e.pmf0[lt][pt])))
                if a["prt_c"][p] == 2 and r <= e.pmf0[lt][pt]:
                    psp += a["psp"][lt][a["prt_seq3r"][p]]
                if e.hdb0[lt][pt] > np.float32(0.1):
                    h3 = (r - e.hdb0[lt][pt]) * e.hdb1[lt][pt]
                    hdb += e.hdb1[lt][pt] * math.exp(-0.5 * h3 * h3)
            hpc += 0.5 * ((h1 - e.hpl0[lt]) / e.hpl1[lt]) ** 2 - e.hpl2[lt]
        kde = 0.0
        for l in range(X.shape[0]):
            sel = a["kde_t"] == a["lig_t"][l]
            if sel.any():
                K = np.stack([a["kde_x"][sel], a["kde_y"][sel], a["kde_z"][sel]], 1)
                kde += np.exp(e.kde2 * ((X[l] - K) ** 2).sum(1)).mean()
        lhm = 0.0
        for r_ in range(a["mcs_x"].shape[0]):
            ok = a["mcs_x"][r_] != np.float32(g.abi.GD_MCS_INVALID)
            T = np.stack([a["mcs_x"][r_][ok], a["mcs_y"][r_][ok], a["mcs_z"][r_][ok]], 1)
            lhm += a["mcs_tcc"][r_] * math.sqrt(((X[ok] - T) ** 2).sum() / ok.sum())
        lhm = math.log(lhm / a["mcs_x"].shape[0])
        dst = np.linalg.norm(a["lig_center"] + m[:3] - a["pocket_center"])
        out.append([vdw, ele, pmf, psp, -hdb / math.sqrt(2 * math.pi), hpc, kde / e.kde3, lhm, dst])
    return np.array(out)


def test_brute_force_tiny(ene, oracle):
    job = g.synth_complex(40, 5, 2, 2, 1, seed=7, ene=ene)
    conf, pose = random_poses(job, 6, seed=4, trans=0.5)
    _, r64, s64, _ = oracle.energy(job, conf, pose, mode=2)
    bf = _brute(job, conf, pose)
    assert rel_err(r64, bf, s64).max() < 1e-6


def test_closed_forms(ene):
    """Properties of the pair terms the restatement must keep."""
    el0, a1, b1 = ene.el0, ene.a1, ene.b1
    assert abs(el0 + a1 + b1 - 1.0) < 1e-6          # g(1) = 1/1
    assert abs(2 * a1 + 3 * b1 + 1.0) < 1e-6        # g'(1) = -1/s^2 at s = 1
    x = 81.0
    poly = 3.5 / 81 * x - 4.5 / 81 ** 2 * x ** 2 + 2.5 / 81 ** 3 * x ** 3 - 0.5 / 81 ** 4 * x ** 4
    assert abs(1 - poly) < 1e-12                    # hpc vanishes at the 9 A cutoff


def test_closed_forms_in_oracle(ene, oracle):
    """One protein point, one ligand atom: pmf at r = pmf0 and at r = 12 A is
    pmf1/2 (the exponent (-r/2 + 6)(r - pmf0) vanishes), hdb at r = hdb0 is
    hdb1; far away vdw/ele/psp/hdb vanish while pmf tends to pmf1."""
    lt, pt = 16, 12          # O.2 vs K1 has an H-bond row in the force field
    assert ene.hdb0[lt][pt] > 0.1
    r0 = float(ene.hdb0[lt][pt])
    base = dict(lig_x=[[0.0]], lig_y=[[0.0]], lig_z=[[0.0]], lig_t=[lt], lig_c=[0.0],
                lig_center=[0, 0, 0], prt_t=[pt], prt_c=[1], prt_cdc=[0], prt_seq3r=[1],
                prt_hpp=[0.0], prt_ele=[0.0], pocket_center=[0, 0, 0],
                psp=np.zeros((24, 2)), kde_t=[lt], kde_x=[0.0], kde_y=[0.0], kde_z=[0.0],
                mcs_x=[[0.0]], mcs_y=[[0.0]], mcs_z=[[0.0]], mcs_tcc=[1.0], minus_beta=[-1.0])
    for r, check in [(r0, "hdb"), (float(ene.pmf0[lt][pt]), "pmf"), (12.0, "pmf")]:
        job = g.HostJob.from_arrays(ene, prt_x=[[r]], prt_y=[[0.0]], prt_z=[[0.0]], **base)
        _, raw, _, _ = oracle.energy(job, np.zeros((1, 2), np.int32), np.zeros((1, 6), np.float32), mode=2)
        if check == "hdb":
            assert abs(raw[0, 4] + ene.hdb1[lt][pt] / math.sqrt(2 * math.pi)) < 1e-6
        else:
            assert abs(raw[0, 2] - ene.pmf1[lt][pt] / 2) < 1e-6
    job = g.HostJob.from_arrays(ene, prt_x=[[500.0]], prt_y=[[0.0]], prt_z=[[0.0]], **base)
    _, raw, _, _ = oracle.energy(job, np.zeros((1, 2), np.int32), np.zeros((1, 6), np.float32), mode=2)
    assert np.all(np.abs(raw[0, [0, 1, 3, 4]]) < 1e-12)
    assert abs(raw[0, 2] - ene.pmf1[lt][pt]) < 1e-12


def test_exchange_detailed_balance(oracle):
    """Two temperatures with frozen energies: the chain alternates between
    'low energy at the low temperature' (swap with p = exp((b0-b1)(E_hi-E_lo)))
    and the reverse (swap always), so the swap rate per attempt is 2p/(1+p)."""
    L = oracle.lib
    L.or_exchange_toy.restype = C.c_double
    L.or_exchange_toy.argtypes = [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                                  C.c_uint64, C.POINTER(C.c_int)]
    for e0, e1 in [(1.0, 0.4), (0.2, 0.9), (0.5, 0.5)]:
        E = (C.c_float * 2)(e0, e1)
        mb = (C.c_float * 2)(-1 / 0.3, -1 / 0.5)
        n = 40000
        rate = L.or_exchange_toy(2, E, mb, n, 12345, None)
        p = min(1.0, math.exp((mb[1] - mb[0]) * (min(e0, e1) - max(e0, e1))))
        expect = 2 * p / (1 + p)
        attempts = n // 2
        assert abs(rate - expect) < 5 * math.sqrt(max(expect * (1 - expect), 1e-9) / attempts) + 1e-9


def test_exchange_ladder_is_permutation(oracle):
    L = oracle.lib
    L.or_exchange_toy.restype = C.c_double
    L.or_exchange_toy.argtypes = [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                                  C.c_uint64, C.POINTER(C.c_int)]
    nt = 8
    E = (C.c_float * nt)(*np.linspace(1.0, -1.0, nt))
    mb = (C.c_float * nt)(*(-1.0 / np.geomspace(0.004, 0.04, nt)))
    tmp = (C.c_int * nt)()
    L.or_exchange_toy(nt, E, mb, 101, 7, tmp)
    assert sorted(tmp[:]) == list(range(nt))
    # strongly ordered energies: after enough rounds the lowest energy sits at T0
    assert tmp[nt - 1] == 0


def test_swap_detailed_balance(oracle):
    """Pose swap between two conformers with frozen energies e[c][x] (conformer
    c at pose x; conformer 0 starts at X, 1 at Y).  Detailed balance of
    u < exp(mb (T_a - E_a) + mb (T_b - E_b)) puts the swapped state at relative
    weight p = exp(mb (e[0][Y] - e[0][X] + e[1][X] - e[1][Y])), so the swap rate
    per attempt is 2 min(1, p) / (1 + p).  A wrong sign, a dropped partner term
    or a decision taken from one member only all change the rate."""
    L = oracle.lib
    L.or_swap_toy.restype = C.c_double
    L.or_swap_toy.argtypes = [C.POINTER(C.c_float), C.c_float, C.c_int, C.c_uint64]
    mb = -1 / 0.4
    for e in [(0.1, 0.5, 0.3, 0.2), (0.5, 0.1, 0.2, 0.6), (0.2, 0.2, 0.7, 0.7), (0.0, 0.9, 0.9, 0.0)]:
        E = (C.c_float * 4)(*e)
        n = 40000
        rate = L.or_swap_toy(E, mb, n, 4321)
        p = math.exp(mb * (e[1] - e[0] + e[2] - e[3]))
        expect = 2 * min(1.0, p) / (1 + p)
        assert abs(rate - expect) < 5 * math.sqrt(max(expect * (1 - expect), 1e-9) / n) + 1e-9, (e, rate)


def swap_attempts(n, rounds):
    """Attempts of each conformer index in `rounds` swap rounds: (c, c+1) for
    c = k mod 2, c+2, ..."""
    cnt = np.zeros(n, np.int64)
    for k in range(rounds):
        for c in range(k & 1, n - 1, 2):
            cnt[c] += 1
            cnt[c + 1] += 1
    return cnt


def same_conformers(job, lig=True, prt=True):
    """Overwrite every ligand (protein) conformer with conformer 0, in place."""
    d = job.desc
    L, P = d.lig_natom, d.prt_npoint
    for on, n, m, arrs in ((lig, d.n_lig, L, (d.lig_x, d.lig_y, d.lig_z)),
                           (prt, d.n_prt, P, (d.prt_x, d.prt_y, d.prt_z))):
        if not on:
            continue
        for a in arrs:
            v = np.ctypeslib.as_array(a, (n * m,))
            v[m:] = np.tile(v[:m], n - 1)


def test_swap_identical_conformers(ene, oracle):
    """With identical conformers a pose swap changes no energy (T_a = E_b and
    T_b = E_a exactly), so the exponent is 0 and every attempt is accepted: the
    swap counts equal the attempt counts of the pairing rule, temperatures stay
    put, and every replica's stored terms are those of its final pose."""
    job = g.synth_complex(300, 16, 4, 3, 2, floor_temp=0.05, ceiling_temp=0.2, seed=4, ene=ene)
    same_conformers(job)
    nl, np_, nt = 4, 3, 2
    mc = g.make_mc(40, steps_per_exchange=5, move_mode="gaussian", seed=3, axes=("lig", "prt"))
    out, _, _ = oracle.dock(job, mc, threads=8)
    r = np.arange(job.n_rep)
    l, p = r % nl, r // (nl * nt)
    assert np.array_equal(out.n_swap_lig, swap_attempts(nl, 8)[l])
    assert np.array_equal(out.n_swap_prt, swap_attempts(np_, 8)[p])
    assert np.array_equal(out.idx_tmp, (r // nl) % nt) and out.n_exchange.sum() == 0
    conf = np.stack([l, p], 1).astype(np.int32)
    terms, _, _, _ = oracle.energy(job, conf, out.movematrix, mode=0)
    assert np.array_equal(terms.astype(np.float32), out.energy)


def test_swap_off_leaves_chain(ene, oracle):
    """Temperature-only exchange (the default, axes 0) and an explicit
    temperature axis give the same chain; conformer axes change it."""
    job = g.synth_complex(300, 16, 3, 2, 3, floor_temp=0.05, ceiling_temp=0.2, seed=6, ene=ene)
    a = g.make_mc(30, steps_per_exchange=5, seed=2)
    a.exchange_axes = 0
    b = g.make_mc(30, steps_per_exchange=5, seed=2, axes=("temp",))
    c = g.make_mc(30, steps_per_exchange=5, seed=2, axes=("temp", "lig", "prt"))
    oa, ob, oc = (oracle.dock(job, m, threads=8)[0] for m in (a, b, c))
    assert np.array_equal(oa.energy, ob.energy) and np.array_equal(oa.idx_tmp, ob.idx_tmp)
    assert oa.n_swap_lig.sum() == 0 and oc.n_swap_lig.sum() > 0 and oc.n_swap_prt.sum() > 0
    assert oc.n_swap_lig.sum() % 2 == 0 and oc.n_swap_prt.sum() % 2 == 0


def test_cpu_dock_entry_point(golden_job, oracle):
    """gd_cpu_dock (the CPU backend's Dock in C-ABI form): the initial pass of the
    golden complex reproduces the published line; a short run equals or_dock in
    the reference's own FP32 expression order."""
    from geauxdock_amd import abi
    from geauxdock_amd.host import ReplicaOut
    L = oracle.lib
    L.gd_cpu_dock.argtypes = [C.c_void_p, C.POINTER(abi.gd_mc_param), C.POINTER(abi.gd_replica_out)]
    mc = g.make_mc(0, move_mode="reference")
    out = ReplicaOut(golden_job.n_rep)
    assert L.gd_cpu_dock(C.byref(golden_job.desc), C.byref(mc), C.byref(out.c)) == 0
    assert np.all(out.is_accept == 1)
    assert " ".join("%.4f" % x for x in out.energy[0]) == " ".join("%.4f" % x for x in GOLDEN)
    mc = g.make_mc(5, move_mode="gaussian", seed=9)
    out = ReplicaOut(golden_job.n_rep)
    assert L.gd_cpu_dock(C.byref(golden_job.desc), C.byref(mc), C.byref(out.c)) == 0
    ref, _, _ = oracle.dock(golden_job, mc, threads=2, literal=True)
    assert np.array_equal(out.energy, ref.energy) and np.array_equal(out.n_accept, ref.n_accept)


@pytest.mark.parametrize("el1", [0.5, 0.1, 1.0])
def test_far_field_threshold(el1):
    """The pair kernel's far-field r^2 threshold (library, FP64 host code): from it
    on the PMF exponent (6 - r/2)(r - pmf0) log2(e) stays <= -26 and el1 r >= 1, so
    the shortcut (PMF term = pmf1, electrostatic factor 1/r) is exact up to FP32
    rounding; and it is tight (2% closer violates one of the two)."""
    log2e = 1.0 / math.log(2.0)
    for pmf0 in np.linspace(0.0, 8.0, 17):
        r = math.sqrt(g.lib().gd_far_r2(float(pmf0), el1))
        for rp in r * (1.0 + np.linspace(0.0, 4.0, 400)):
            assert log2e * (6.0 - rp / 2.0) * (rp - pmf0) <= -26.0 + 1e-9, (pmf0, rp)
            assert el1 * rp >= 1.0
        rs = 0.98 * r
        assert log2e * (6.0 - rs / 2.0) * (rs - pmf0) > -26.0 or el1 * rs < 1.0


def test_near_branch_threshold(ene):
    """The row threshold of the pair loop's near branch (library, FP64 host code),
    for every (ligand type, point type) row of the published force field: it is
    at least the far-field threshold, and from it on the terms the branch also
    holds are exactly zero in the kernel's FP32 arithmetic (hydrophobic r^2 <= 81,
    psp sqrtf(r^2) <= pmf0, H-bond c exp2(-h^2) with h = fma(r, c, -hdb0 c)), so
    a wave whose lanes are all beyond it may skip the branch bit-exactly."""
    f32 = np.float32
    k = math.sqrt(0.5 / math.log(2.0))
    pmf0 = np.array(ene.pmf0, np.float32)
    hdb0 = np.array(ene.hdb0, np.float32)
    hdb1 = np.array(ene.hdb1, np.float32)
    n_hb = 0
    for lt in range(pmf0.shape[0]):
        for pt in range(pmf0.shape[1]):
            p0, h0, h1 = float(pmf0[lt, pt]), float(hdb0[lt, pt]), float(hdb1[lt, pt])
            if not (np.isfinite(p0) and np.isfinite(h0) and np.isfinite(h1)):
                continue
            r2n = g.lib().gd_near_r2(p0, ene.el1, h0, h1)
            assert r2n >= g.lib().gd_far_r2(p0, ene.el1)
            r2 = (f32(r2n) * (1.0 + np.linspace(0.0, 8.0, 2001))).astype(f32)
            r2 = np.concatenate([[f32(r2n)], r2])
            assert np.all(r2 > f32(81.0))
            assert np.all(np.sqrt(r2).astype(f32) > f32(p0))          # no psp contact
            if f32(h0) > f32(0.1):
                n_hb += 1
                c = f32(h1 * k)
                n0 = f32(-float(h0) * float(c))
                rr = (r2 * (f32(1.0) / np.sqrt(r2).astype(f32))).astype(f32)
                h = (rr.astype(np.float64) * float(c) + float(n0)).astype(f32)
                assert np.all(np.exp2(-(h * h)).astype(f32) == 0.0), (lt, pt)
    assert n_hb > 0
