"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the reference CPU backend (oracle/_ref/libgdref.so).  Test code only.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from geauxdock_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libgdref.so")
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE):
            build_oracle()
        self.lib = C.CDLL(ORACLE)
        L = self.lib
        L.or_energy.argtypes = [C.c_void_p, C.c_int, _ip, _fp, _fp, C.c_int, _dp, _dp, _dp, _fp, C.c_int]
        L.or_dock.argtypes = [C.c_void_p, C.POINTER(abi.gd_mc_param), C.c_int, C.c_int, C.c_int,
                              C.POINTER(abi.gd_replica_out), _fp, _dp, _fp, C.c_int, _ip, C.c_int]
        L.or_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_uint32)]
        L.or_step_draws.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
        L.or_exp_underflow_threshold.restype = C.c_float

    def energy(self, job, conf, pose, mode=2, coords=None, threads=0, want_coords=False):
        """mode 0 = FP32 (GPU rounding), 1 = FP32 reference-literal, 2 = FP64."""
        conf = np.ascontiguousarray(conf, np.int32)
        pose = np.ascontiguousarray(pose, np.float32)
        n = conf.shape[0]
        terms = np.zeros((n, 10))
        raw = np.zeros((n, 9))
        scale = np.zeros((n, 9))
        cin = None if coords is None else np.ascontiguousarray(coords, np.float32)
        cout = np.zeros((n, job.desc.lig_natom, 3), np.float32) if want_coords else None
        rc = self.lib.or_energy(C.byref(job.desc), n, conf.ctypes.data_as(_ip), pose.ctypes.data_as(_fp),
                                None if cin is None else cin.ctypes.data_as(_fp), mode,
                                terms.ctypes.data_as(_dp), raw.ctypes.data_as(_dp),
                                scale.ctypes.data_as(_dp),
                                None if cout is None else cout.ctypes.data_as(_fp), threads)
        assert rc == 0
        return terms, raw, scale, cout

    def dock(self, job, mc, threads=0, rep_lo=0, rep_hi=0, trace=False, records=False, literal=False):
        """records=True adds (rec[n_rep][steps_total][17], count[n_rep]): the
        accept records of the whole run.  literal=True evaluates in the
        reference's own FP32 expression order (else the GPU rounding)."""
        from geauxdock_amd.host import ReplicaOut
        n = job.n_rep
        out = ReplicaOut(n)
        tr = np.zeros((n, max(1, mc.steps_total)), np.float32) if trace else None
        cap = max(1, mc.steps_total)
        rec = np.zeros((n, cap, 17), np.float32) if records else None
        cnt = np.zeros(n, np.int32) if records else None
        ms = C.c_double()
        rc = self.lib.or_dock(C.byref(job.desc), C.byref(mc), threads, rep_lo, rep_hi, C.byref(out.c),
                              None if tr is None else tr.ctypes.data_as(_fp), C.byref(ms),
                              None if rec is None else rec.ctypes.data_as(_fp), cap,
                              None if cnt is None else cnt.ctypes.data_as(_ip), 1 if literal else 0)
        assert rc == 0
        if records:
            return out, tr, ms.value, (rec, cnt)
        return out, tr, ms.value

    def philox(self, ctr, key, rounds=10):
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        self.lib.or_philox(c, k, rounds, o)
        return list(o)

    def draws(self, seed, job, rep, step):
        w = (C.c_uint32 * 8)()
        self.lib.or_step_draws(seed, job, rep, step, w)
        return list(w)


def random_poses(job, n, seed, trans=1.5, rot=np.pi):
    rng = np.random.default_rng(seed)
    d = job.desc
    conf = np.stack([rng.integers(0, d.n_lig, n), rng.integers(0, d.n_prt, n)], 1).astype(np.int32)
    pose = np.concatenate([rng.normal(0, trans, (n, 3)), rng.uniform(-rot, rot, (n, 3))], 1)
    return conf, pose.astype(np.float32)


def rel_err(a, b, scale):
    """|a - b| / max(|b|, scale): the per-term error measure of the tests."""
    return np.abs(np.asarray(a, np.float64) - b) / np.maximum(np.abs(b), np.maximum(scale, 1e-30))
