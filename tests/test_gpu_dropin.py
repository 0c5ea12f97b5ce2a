"""The reference backend entry points end to end: tests/dropin/dropin_main runs
the reference GPU driver's call sequence (src/main_gpu/geauxdock.C:59-104:
SetDevice, DeviceAlloc, InitCurand, SetParaT, CopyH2D, Dock, DeviceFree)
against the library through include/geauxdock_dropin.h, on the fixture
complex filled into the reference's own Complex layout."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN_DATA

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "dropin", "dropin_main")
PARAMS = os.path.join(ROOT, "geauxdock_amd", "data", "parameters")
GOLDEN = "0.6434 -0.0367 -0.2079 -0.1845 0.8521 -0.8880 0.0523 0.1744 -1.0000 0.7735"


def _run(*args):
    if not os.path.exists(BIN):
        pytest.fail("tests/dropin/dropin_main is not built (run __graft_entry__.build())")
    env = dict(os.environ, GD_SEED="7")
    r = subprocess.run([BIN, GOLDEN_DATA, PARAMS] + [str(a) for a in args], capture_output=True,
                       text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def _finals(lines):
    rows = [l.split() for l in lines if l.startswith("final ")]
    return np.array([[float(x) for x in r[2:14]] for r in rows]), np.array([int(r[15]) for r in rows])


def test_dropin_dock_reports_reference_line():
    """Replica 0's first record of the last dump is its initial state: the
    printed terms equal the reference's published line (its "Ref Result")."""
    lines = _run(10)
    rec0 = [l for l in lines if l.startswith("  0 0 ") and "Ref Result" not in l]
    assert rec0, lines[:20]
    assert " ".join(rec0[0].split()[2:]) == GOLDEN
    assert any(l.startswith("Benchmark,") for l in lines)
    finals, nrec = _finals(lines)
    assert len(finals) == 51 and np.all(np.isfinite(finals))
    assert np.all((nrec >= 1) & (nrec <= 10))


def test_dropin_dock_with_exchange():
    """Four temperatures with exchange every 5 steps: each (ligand, protein)
    ladder still holds every temperature once, some swaps happened, and the run
    is reproducible for a fixed seed (GD_SEED)."""
    a = _run(40, 5, 4)
    b = _run(40, 5, 4)
    fa, _ = _finals(a)
    fb, _ = _finals(b)
    assert np.array_equal(fa, fb)
    nl, np_, nt = [int(x) for x in [l for l in a if l.startswith("sizes ")][0].split()[1:]]
    assert nt == 4 and len(fa) == nl * np_ * nt
    tmp = fa[:, 0].astype(int)
    moved = 0
    for p in range(np_):
        for l in range(nl):
            reps = [nt * nl * p + nl * t + l for t in range(nt)]  # util.C:380
            assert sorted(tmp[reps]) == list(range(nt))
            moved += sum(tmp[r] != t for t, r in enumerate(reps))
    assert moved > 0


def test_dropin_exchange_off_unless_requested():
    """The frontend never writes McPara.steps_per_exchange; the test driver
    leaves a leftover value (3) in it.  Without GD_EXCHANGE_STEPS the drop-in
    Dock must not exchange, as the reference GPU Dock never does: every
    replica keeps its own temperature."""
    lines = _run(40, 0, 4)
    fa, _ = _finals(lines)
    nl, np_, nt = [int(x) for x in [l for l in lines if l.startswith("sizes ")][0].split()[1:]]
    tmp = fa[:, 0].astype(int)
    for p in range(np_):
        for t in range(nt):
            for l in range(nl):
                assert tmp[nt * nl * p + nl * t + l] == t


def test_dropin_records_every_dump():
    """450 steps in dumps of 200: the last dump (steps 400..449) is what rh
    holds, as after the reference's per-dump copies; replica 0's first record
    of it is at a step >= 400 or it has none."""
    lines = _run(450)
    rec0 = [l for l in lines if l.startswith("  0 ") and "Ref Result" not in l]
    _, nrec = _finals(lines)
    assert np.all((nrec >= 0) & (nrec <= 50))
    if nrec[0] > 0:
        assert int(rec0[0].split()[1]) >= 400
