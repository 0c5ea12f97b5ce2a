"""The move generator against the mathematics (CPU).

The reference move is m += scale * U[0,1) (backend_cpu_mic/
kernel_cpp_montecarlo_soa.C:242-250); the benchmark configurations ask for
N(0, 0.2 A) translations and N(0, 3 deg) rotations, which this library draws
by Box-Muller from the same Philox words.  Both the library's draw path
(gd_move_draws, the kernels' source compiled for the host) and the oracle's
own restatement are checked here against the distributions they must follow,
so the GPU-versus-oracle chain tests, which share the formula, inherit a pin
that does not come from retyping it:
  - Gaussian mode: per component mean 0 and variance scale^2 (5 sigma bounds
    over 10^6 increments), a Kolmogorov-Smirnov test against N(0, scale), no
    correlation inside a Box-Muller pair, finite draws at the extreme words,
    and the closed forms at words that fix u1 and u2;
  - reference mode: every increment in [0, scale), mean scale / 2, variance
    scale^2 / 12, KS against U(0, scale).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats

import geauxdock_amd as g
from geauxdock_amd import abi

_fp = C.POINTER(C.c_float)
SCALE = np.array([0.2, 0.2, 0.2, math.radians(3.0), math.radians(3.0), math.radians(3.0)],
                 np.float32)
N_REP, N_STEPS = 4096, 50          # 204,800 replica-steps = 1,228,800 increments


def lib_draws(mode, scale=SCALE, seed=12345, job=3, n_rep=N_REP, n_steps=N_STEPS, step0=0,
              device=0):
    inc = np.zeros((n_steps, n_rep, 6), np.float32)
    sc = np.ascontiguousarray(scale, np.float32)
    g.check(g.lib().gd_move_draws(seed, job, mode, sc.ctypes.data_as(_fp), n_rep, step0, n_steps,
                                  device, inc.ctypes.data_as(_fp)), "gd_move_draws")
    return inc.reshape(-1, 6)


def oracle_draws(oracle, mode, scale=SCALE, seed=12345, job=3, n_rep=N_REP, n_steps=N_STEPS):
    inc = np.zeros((n_steps, n_rep, 6), np.float32)
    sc = np.ascontiguousarray(scale, np.float32)
    f = oracle.lib.or_move_draws
    f.argtypes = [C.c_uint64, C.c_uint32, C.c_int, _fp, C.c_int, C.c_uint32, C.c_int, _fp]
    f(seed, job, mode, sc.ctypes.data_as(_fp), n_rep, 0, n_steps, inc.ctypes.data_as(_fp))
    return inc.reshape(-1, 6)


def check_gaussian(inc, scale=SCALE):
    n = inc.shape[0]
    x = inc.astype(np.float64)
    assert np.all(np.isfinite(x))
    for k in range(6):
        s = float(scale[k])
        m, v = x[:, k].mean(), x[:, k].var()
        assert abs(m) < 5 * s / math.sqrt(n), (k, m)
        assert abs(v - s * s) < 5 * s * s * math.sqrt(2.0 / n), (k, v, s * s)
        p = stats.kstest(x[:, k] / s, "norm").pvalue
        assert p > 1e-4, (k, p)
    for k in range(0, 6, 2):   # the two members of one Box-Muller pair
        rho = np.corrcoef(x[:, k], x[:, k + 1])[0, 1]
        assert abs(rho) < 5 / math.sqrt(n), (k, rho)
    # the squared radius of a pair, scaled, is chi-square with 2 degrees of freedom
    r2 = (x[:, 0] / scale[0]) ** 2 + (x[:, 1] / scale[1]) ** 2
    assert stats.kstest(r2, "chi2", args=(2,)).pvalue > 1e-4


def check_uniform(inc, scale=SCALE):
    n = inc.shape[0]
    x = inc.astype(np.float64)
    for k in range(6):
        s = float(scale[k])
        assert x[:, k].min() >= 0.0 and x[:, k].max() < s, k
        assert abs(x[:, k].mean() - s / 2) < 5 * s / math.sqrt(12 * n)
        assert abs(x[:, k].var() - s * s / 12) < 5 * s * s * math.sqrt(1.0 / (80 * n)) * 2
        assert stats.kstest(x[:, k] / s, "uniform").pvalue > 1e-4, k


def test_library_gaussian_moves():
    check_gaussian(lib_draws(abi.GD_MOVE_GAUSSIAN))


def test_library_reference_uniform_moves():
    check_uniform(lib_draws(abi.GD_MOVE_REF_UNIFORM))


def test_oracle_gaussian_moves(oracle):
    inc = oracle_draws(oracle, abi.GD_MOVE_GAUSSIAN)
    check_gaussian(inc)
    # the oracle and the library draw the same numbers (same words, same formula)
    ref = lib_draws(abi.GD_MOVE_GAUSSIAN)
    ulp = np.spacing(np.maximum(np.abs(ref), 1e-6).astype(np.float32))
    assert np.all(np.abs(inc - ref) <= 2 * ulp)


def test_oracle_reference_uniform_moves(oracle):
    inc = oracle_draws(oracle, abi.GD_MOVE_REF_UNIFORM)
    check_uniform(inc)
    assert np.array_equal(inc, lib_draws(abi.GD_MOVE_REF_UNIFORM))


def test_gaussian_closed_forms():
    """Words that fix the uniforms: u = (w >> 8) 2^-24.  u1 = 0 gives radius 0;
    u1 = 1 - 2^-24 gives the largest radius sqrt(48 ln 2) (finite: the draw uses
    1 - u1); u2 = 0 and 1/4 put the pair on the axes; u2 = 1/8 on the diagonal."""
    top = 0xFFFFFF00
    rmax = math.sqrt(48 * math.log(2.0))
    cases = [  # (w1, w2) -> expected (even, odd) component for unit scale
        (0, 0, 0.0, 0.0),
        (top, 0, rmax, 0.0),
        (top, 0x40000000, 0.0, rmax),
        (top, 0x20000000, rmax / math.sqrt(2), rmax / math.sqrt(2)),
        (0x80000000, 0x80000000, -math.sqrt(2 * math.log(2.0)), 0.0),
    ]
    w = np.zeros((len(cases), 8), np.uint32)
    for i, (a, b, _, _) in enumerate(cases):
        w[i, 0:6] = [a, b, a, b, a, b]
    scale = np.array([1, 1, 2, 2, 0.5, 0.5], np.float32)
    inc = np.zeros((len(cases), 6), np.float32)
    g.check(g.lib().gd_move_incr_words(abi.GD_MOVE_GAUSSIAN, scale.ctypes.data_as(_fp), len(cases),
                                       w.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       inc.ctypes.data_as(_fp)), "gd_move_incr_words")
    for i, (_, _, e0, e1) in enumerate(cases):
        for k in range(3):
            s = scale[2 * k]
            assert abs(inc[i, 2 * k] - s * e0) <= 1e-5 * s * rmax, (i, k, inc[i])
            assert abs(inc[i, 2 * k + 1] - s * e1) <= 1e-5 * s * rmax, (i, k, inc[i])
    # reference mode: the word's top 24 bits over 2^24, times the scale
    g.check(g.lib().gd_move_incr_words(abi.GD_MOVE_REF_UNIFORM, scale.ctypes.data_as(_fp), len(cases),
                                       w.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       inc.ctypes.data_as(_fp)), "gd_move_incr_words")
    u = (w[:, :6] >> 8).astype(np.float64) / 2 ** 24
    assert np.array_equal(inc, (scale * u).astype(np.float32))


def test_draws_depend_on_every_counter_word():
    base = lib_draws(abi.GD_MOVE_GAUSSIAN, n_rep=64, n_steps=4)
    for kw in [dict(seed=12346), dict(job=4), dict(step0=1)]:
        other = lib_draws(abi.GD_MOVE_GAUSSIAN, n_rep=64, n_steps=4, **kw)
        assert not np.any(np.all(other == base, axis=1)), kw
