"""GPU parity of the energy path (gd_energy_eval: transform + pair kernel +
finish kernel) against the CPU oracle.

Tolerance (the per-term bar): every raw energy term of every pose within
1e-5 relative of the FP64 oracle, relative to max(|value|, sum of |pair
contributions|) -- the sum of magnitudes is the scale at which a
mixed-sign sum of 10^5 pair terms can be resolved in FP32 at all.  The
oracle is replayed at the GPU's own transformed coordinates so that the
discontinuous psp cutoff decides on identical distances; the transform is
checked separately.
"""
import os

import numpy as np
import pytest

import geauxdock_amd as g
from tests.oracle_ffi import random_poses, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5
# pose counts of the random-pose parity tests; GD_PARITY_POSES raises them for
# a longer sweep (the defaults keep the suite to about a minute)
N_POSES = int(os.environ.get("GD_PARITY_POSES", "0"))
GOLDEN = np.array([0.6434, -0.0367, -0.2079, -0.1845, 0.8521, -0.8880, 0.0523, 0.1744, -1.0000,
                   0.7735])


@pytest.fixture(scope="module")
def golden_dev(golden_job):
    return g.DeviceJob(golden_job)


# Transform bar: each GPU coordinate x = R a + (m + center) within
# TRANSFORM_ULPS units in the last place of its operands' scale
# max(|x|, |m + center|, conformer radius) -- an FP32 sum rounds relative to
# its operands, so a coordinate near 0 A built from 40 A terms is held to the
# ulp of 40 A, not of its own value.
TRANSFORM_ULPS = 4


def _transform_scale(job, pose, x):
    a = job.arrays()
    rad = float(np.sqrt(a["lig_x"] ** 2 + a["lig_y"] ** 2 + a["lig_z"] ** 2).max())
    t = np.abs(pose[:, None, :3].astype(np.float64) + np.array(job.desc.lig_center[:3], np.float64))
    s = np.maximum(np.maximum(np.abs(x), t), rad).astype(np.float32)
    return np.spacing(s).astype(np.float64)


def _parity(job, dev, oracle, conf, pose, label, check_transform=True):
    terms, raw, coords = dev.eval(conf, pose, want_coords=True)
    if check_transform:
        # transform parity: GPU sincos/FMA against the FP64 transform, in ulps
        _, _, _, o_c = oracle.energy(job, conf, pose, mode=2, want_coords=True)
        t_ulps = np.abs(coords.astype(np.float64) - o_c) / _transform_scale(job, pose, o_c)
        print("\n[%s] transform: max %.2f ulp" % (label, t_ulps.max()))
        assert t_ulps.max() <= TRANSFORM_ULPS, "transform differs by %.1f ulp" % t_ulps.max()
    # energy parity at the GPU coordinates
    r_t, r_raw, r_scale, _ = oracle.energy(job, conf, pose, mode=2, coords=coords)
    err = rel_err(raw, r_raw, r_scale)
    worst = err.max(0)
    # the plain relative error |gpu - fp64| / |fp64|, reported beside the bar's
    # measure; it is the bar itself for the terms whose contributions share a
    # sign (H-bond, KDE: scale == |value|)
    plain = (np.abs(raw - r_raw) / np.maximum(np.abs(r_raw), 1e-30)).max(0)
    print("[%s] max rel err per raw term vs FP64 (scaled | plain):" % label,
          " ".join("%s=%.2e|%.2e" % (n, e, q) for n, e, q in zip(g.abi.TERMS[:9], worst, plain)))
    assert np.all(np.isfinite(raw)), "non-finite energy"
    assert worst.max() <= TOL, worst
    # the normalised terms and the total follow the same bar on their own scale
    d = job.ene
    a = np.array(d.a_para[:9], np.float64)
    L = job.desc.lig_natom
    scale_n = np.abs(a) * r_scale / np.where(np.arange(9) < 7, L, 1)
    err_n = np.abs(terms[:, :9] - r_t[:, :9]) / np.maximum(np.abs(r_t[:, :9]), scale_n)
    assert err_n.max() <= TOL, err_n.max(0)
    w = np.array(d.w[:9], np.float64)
    tot_scale = (np.abs(w) * np.maximum(np.abs(r_t[:, :9]), scale_n)).sum(1)
    assert (np.abs(terms[:, 9] - r_t[:, 9]) / tot_scale).max() <= TOL
    return terms, raw


def test_golden_line(golden_job, golden_dev, oracle):
    conf = np.zeros((1, 2), np.int32)
    pose = np.zeros((1, 6), np.float32)
    terms, raw = _parity(golden_job, golden_dev, oracle, conf, pose, "golden")
    # the published line is printed with %.4f (doc/instructions.txt:186)
    assert np.all(np.abs(terms[0] - GOLDEN) <= 0.5e-4 + 1e-6), terms[0]
    assert " ".join("%.4f" % x for x in terms[0]) == " ".join("%.4f" % x for x in GOLDEN)


def test_golden_random_poses(golden_job, golden_dev, oracle):
    n = max(512, N_POSES)
    conf, pose = random_poses(golden_job, n, seed=11)
    _parity(golden_job, golden_dev, oracle, conf, pose, "1a07C1 x%d" % n)


def test_golden_all_conformers_native(golden_job, golden_dev, oracle):
    n = golden_job.desc.n_lig
    conf = np.stack([np.arange(n), np.zeros(n)], 1).astype(np.int32)
    _parity(golden_job, golden_dev, oracle, conf, np.zeros((n, 6), np.float32), "conformers")


@pytest.mark.parametrize("P,L,nl,np_,seed", [(1200, 24, 5, 5, 1), (3000, 40, 20, 11, 2),
                                             (9000, 80, 3, 2, 4), (1500, 12, 4, 3, 5)])
def test_synthetic_parity(ene, oracle, P, L, nl, np_, seed):
    job = g.synth_complex(P, L, nl, np_, 2, seed=seed, ene=ene)
    dev = g.DeviceJob(job)
    conf, pose = random_poses(job, max(96, N_POSES), seed=seed + 100, trans=1.0)
    _parity(job, dev, oracle, conf, pose, "synthetic P=%d L=%d" % (P, L))


@pytest.mark.parametrize("P,L", [(8, 2), (37, 3), (64, 5), (130, 7)])
def test_small_ragged_parity(ene, oracle, P, L):
    """Ragged shapes: ligand chunks padded with empty atoms (L not a multiple of
    the chunk width), proteins smaller than one work group's eight tiles, and
    one-point tiles, which leave whole tiles without points, KDE anchors or
    MCS rows (the tile count is rounded up to a multiple of eight)."""
    job = g.synth_complex(P, L, 3, 2, 2, seed=P + L, ene=ene)
    dev = g.DeviceJob(job)
    conf, pose = random_poses(job, 96, seed=P, trans=1.0)
    first = None
    for tiling in [None, (4, 1), (8, 3), (16, 5)]:
        if tiling:
            dev.set_tiling(*tiling)
            lc, tp = dev.tiling()
            assert (lc, tp) == (tiling[0], min(tiling[1], P)), (lc, tp)
        _parity(job, dev, oracle, conf, pose, "P=%d L=%d tiling %s" % (P, L, tiling),
                check_transform=tiling is None)
        terms, _, _ = dev.eval(conf, pose)
        if first is None:
            first = terms
        else:   # only the summation order changes with the tiling
            assert np.abs(terms - first).max() < 1e-4


def test_eval_of_no_poses(golden_dev):
    terms, raw, _ = golden_dev.eval(np.zeros((0, 2), np.int32), np.zeros((0, 6), np.float32))
    assert terms.shape == (0, 10) and raw.shape == (0, 9)


def test_far_away_and_edges(golden_job, golden_dev, oracle):
    # far from the protein: pair terms vanish, dst dominates; exact zero pose
    conf = np.zeros((3, 2), np.int32)
    pose = np.array([[60, 0, 0, 0, 0, 0], [0, -45, 20, 1, 2, 3], [0, 0, 0, 0, 0, 0]], np.float32)
    _parity(golden_job, golden_dev, oracle, conf, pose, "edges")


def test_eval_more_poses_than_replicas(golden_job, golden_dev, oracle):
    # 51 replicas pad to 64 lanes; 150 poses run in three chunks
    conf, pose = random_poses(golden_job, 150, seed=3)
    t1, _, _ = golden_dev.eval(conf, pose)
    t2, _, _ = golden_dev.eval(conf[:40], pose[:40])
    assert np.array_equal(t1[:40], t2), "a pose's energy must not depend on its batch"


def test_tiling_invariance(ene, oracle):
    job = g.synth_complex(2000, 30, 4, 3, 2, seed=9, ene=ene)
    dev = g.DeviceJob(job)
    conf, pose = random_poses(job, 64, seed=5)
    ref = None
    for lc, pt in [(4, 64), (8, 125), (16, 500), (8, 2000)]:
        dev.set_tiling(lc, pt)
        assert dev.tiling() == (lc, pt)
        t, raw, _ = dev.eval(conf, pose)
        r_t, r_raw, r_scale, _ = oracle.energy(job, conf, pose, mode=2,
                                               coords=dev.eval(conf, pose, True)[2])
        assert rel_err(raw, r_raw, r_scale).max() <= TOL
        if ref is not None:
            assert np.abs(t - ref).max() < 1e-4
        ref = t


@pytest.mark.parametrize("el0,el1", [(2.0, 0.1), (1.5, 0.1)])
def test_electrostatics_short_range(ene, oracle, el0, el1):
    """el1 = 0.1 puts every pair under 10 A on the polynomial branch of the
    electrostatic factor (and out of the far-field shortcut).  el0 = 2 is the
    reference value, continuous at s = 1; el0 = 1.5 makes the branch jump there."""
    import ctypes as C
    from geauxdock_amd import abi
    e = abi.gd_enepara()
    C.memmove(C.byref(e), C.byref(ene), C.sizeof(abi.gd_enepara))
    e.el0, e.el1 = el0, el1
    job = g.synth_complex(1500, 24, 3, 2, 2, seed=17, ene=e)
    dev = g.DeviceJob(job)
    conf, pose = random_poses(job, 128, seed=23, trans=1.5)
    _parity(job, dev, oracle, conf, pose, "el0=%g el1=%g" % (el0, el1))


@pytest.mark.parametrize("sigma", [4.0, 6.0, 8.0])
def test_golden_displaced_poses(golden_job, golden_dev, oracle, sigma):
    """Poses displaced sigma A (per axis) from the pocket: the band where the
    H-bond and KDE sums are carried by contributions in the tails of their
    Gaussians (at 8 A many KDE sums are near FP32 underflow)."""
    conf, pose = random_poses(golden_job, max(1000, N_POSES), seed=int(31 + sigma), trans=sigma)
    _parity(golden_job, golden_dev, oracle, conf, pose, "1a07C1 sigma=%g" % sigma)


@pytest.fixture(scope="module")
def c2_job(ene):
    # the 1,760-replica benchmark complex: P = 3,000 points x 11, L = 40 atoms x 20, 8 temperatures
    return g.synth_complex(3000, 40, 20, 11, 8, floor_temp=0.004, ceiling_temp=0.04, seed=2, ene=ene)


@pytest.mark.parametrize("sigma", [4.0, 6.0, 8.0])
def test_c2_displaced_poses(c2_job, oracle, sigma):
    dev = g.DeviceJob(c2_job)
    conf, pose = random_poses(c2_job, max(128, N_POSES // 8), seed=int(41 + sigma), trans=sigma)
    _parity(c2_job, dev, oracle, conf, pose, "C2 sigma=%g" % sigma)


def test_c2_trajectory_poses(c2_job, oracle):
    """The poses an MC run visits: 2,000 Gaussian-move steps of every replica
    with exchange on all three axes, then the final accepted pose of 256
    replicas spread over the ladder.  The terms the run stored for those poses
    must equal a fresh evaluation bit for bit, and both meet the FP64 bar.
    Replicas whose forced initial accept underflowed (initial total above
    104 T, the reference's expf quirk: the four coldest temperatures here)
    never hold a state, so they are left out."""
    dev = g.DeviceJob(c2_job)
    mc = g.make_mc(2000, steps_per_exchange=100, move_mode="gaussian", seed=7,
                   axes=("temp", "lig", "prt"))
    out, _ = dev.dock(mc)
    d = c2_job.desc
    live = np.flatnonzero((out.n_accept > 0) | (out.n_swap_lig > 0) | (out.n_swap_prt > 0))
    assert live.size >= 256
    reps = live[np.linspace(0, live.size - 1, 256).astype(np.int64)]
    conf = np.stack([reps % d.n_lig, reps // (d.n_tmp * d.n_lig)], 1).astype(np.int32)
    pose = np.ascontiguousarray(out.movematrix[reps], np.float32)
    moved = np.linalg.norm(pose[:, :3], axis=1)
    print("\n[C2 trajectory] displacement from the pocket: median %.2f A, max %.2f A"
          % (np.median(moved), moved.max()))
    terms, _ = _parity(c2_job, dev, oracle, conf, pose, "C2 after 2000 MC steps")
    assert np.array_equal(terms, out.energy[reps]), "stored terms differ from a fresh evaluation"
