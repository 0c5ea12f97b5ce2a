"""The tiling table's lookup keys (CPU; gd_tiling_lookup touches no device).

One-job tilings are keyed on the ligand's 4-atom chunk count ceil(L / 4) and on
octaves of P and R; shared-launch tilings on octaves of L and P.  Every entry
of the compiled table must be what the library returns for a job inside its
bucket, and jobs either side of a bucket edge must get their own bucket's
entry."""
import ctypes as C
import math
import os

import pytest

import geauxdock_amd as g
from geauxdock_amd import abi, tune

TABLE = os.path.join(os.path.dirname(g.__file__), "csrc", "gd_tiling_table.h")


def lookup(L, P, R, shared):
    t = abi.gd_tiling()
    g.check(g.lib().gd_tiling_lookup(L, P, R, int(shared), C.byref(t)), "gd_tiling_lookup")
    return t.atoms_per_lane, t.tile_points, t.waves_per_cu


def first_with(key_fn, key, lo=1, hi=4096):
    return next(x for x in range(lo, hi) if key_fn(x) == key)


def test_chunk_key():
    for L in range(1, 500):
        assert tune.chunk_key(L) == math.ceil(L / 4)


@pytest.mark.parametrize("mode", ["GD_TUNE_PLAIN", "GD_TUNE_SHARED"])
def test_every_entry_is_what_the_library_returns(mode):
    entries = tune.read_table(TABLE, (mode,))
    assert entries
    for _, kl, lgP, lgR, lc, value in entries:
        P = 2 ** lgP
        if mode == "GD_TUNE_PLAIN":
            R = 2 ** lgR
            for L in (first_with(tune.chunk_key, kl), max(x for x in range(1, 4096) if tune.chunk_key(x) == kl)):
                got = lookup(L, P, R, False)
                assert got[:2] == (lc, -(-P // min(value, P))), (kl, lgP, lgR, L)
        else:
            for L in (2 ** kl, 2 ** (kl + 1) - 1):
                assert lookup(L, P, 1760, True)[:2] == (4, min(P, value)), (kl, lgP, L)


def test_bucket_edges_pick_their_own_entries():
    plain = {(e[1], e[2], e[3]): e for e in tune.read_table(TABLE, ("GD_TUNE_PLAIN",))}
    for L0, L1 in [(40, 41), (44, 45), (12, 13), (79, 80)]:
        for P, R in [(3000, 1760), (6000, 440)]:
            a, b = (plain.get((tune.chunk_key(L), tune.lg2(P), tune.lg2(R))) for L in (L0, L1))
            if a is None or b is None:
                continue
            assert lookup(L0, P, R, False)[:2] == (a[4], -(-P // min(a[5], P)))
            assert lookup(L1, P, R, False)[:2] == (b[4], -(-P // min(b[5], P)))


def test_table_round_trip(tmp_path):
    """tune.py writes a header the reader (and the C table) take back unchanged."""
    entries = tune.read_table(TABLE, ("GD_TUNE_PLAIN", "GD_TUNE_SHARED", "GD_TUNE_WAVES"))
    out = tmp_path / "t.h"
    tune.write_table(entries, str(out), "round trip")
    assert tune.read_table(str(out), ("GD_TUNE_PLAIN", "GD_TUNE_SHARED", "GD_TUNE_WAVES")) == entries
    text = open(out).read()
    assert "kTunedTable[]" in text and "// round trip" in text
    # one entry per key of each mode
    for mode in ("GD_TUNE_PLAIN", "GD_TUNE_SHARED"):
        keys = [e[1:4] for e in entries if e[0] == mode]
        assert len(keys) == len(set(keys)), mode
