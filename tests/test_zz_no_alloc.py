"""gd_energy_eval_noalloc (no allocation; kernels' scratch in a static device
array) against gd_energy_eval through the C ABI: the same bits, on two calls
in a row, the second with fewer poses, on a job of one 64-lane block and on a
job of 28 blocks."""
import ctypes as C

import numpy as np
import pytest

import geauxdock_amd as g
from tests.oracle_ffi import random_poses

pytestmark = pytest.mark.gpu

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _noalloc(dev, conf, pose):
    conf = np.ascontiguousarray(conf, np.int32)
    pose = np.ascontiguousarray(pose, np.float32)
    n = conf.shape[0]
    terms = np.full((n, 10), np.nan, np.float32)
    raw = np.full((n, 9), np.nan, np.float32)
    g.check(g.lib().gd_energy_eval_noalloc(dev.h, n, conf.ctypes.data_as(_ip), pose.ctypes.data_as(_fp),
                                           terms.ctypes.data_as(_fp), raw.ctypes.data_as(_fp)),
            "gd_energy_eval_noalloc")
    return terms, raw


@pytest.mark.parametrize("shape,n1,n2", [((1200, 24, 5, 5, 2, 1), 150, 40),
                                          ((3000, 40, 20, 11, 8, 2), 300, 70)])
def test_noalloc_matches_eval(ene, shape, n1, n2):
    P, L, nl, np_, nt, seed = shape
    job = g.synth_complex(P, L, nl, np_, nt, seed=seed, ene=ene)
    dev = g.DeviceJob(job)
    for n, s in [(n1, 11), (n2, 12)]:
        conf, pose = random_poses(job, n, seed=s, trans=2.0)
        t_ref, r_ref, _ = dev.eval(conf, pose)
        t, r = _noalloc(dev, conf, pose)
        assert np.array_equal(t, t_ref), np.abs(t - t_ref).max()
        assert np.array_equal(r, r_ref), np.abs(r - r_ref).max()


def test_noalloc_leaves_the_job_alone(ene):
    """The job's replica state and a later MC run are unaffected."""
    job = g.synth_complex(1200, 24, 5, 5, 2, seed=1, ene=ene)
    dev = g.DeviceJob(job)
    mc = g.make_mc(20, move_mode="gaussian", seed=4)
    a, _ = dev.dock(mc)
    conf, pose = random_poses(job, 100, seed=3)
    _noalloc(dev, conf, pose)
    after = dev.replicas()
    assert np.array_equal(after.energy, a.energy) and np.array_equal(after.movematrix, a.movematrix)
    b, _ = dev.dock(mc)
    assert np.array_equal(a.energy, b.energy)
