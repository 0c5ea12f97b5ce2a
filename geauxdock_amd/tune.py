"""Tiling autotuner: sweeps on an MI355X, written out as the tiling table.

    python -m geauxdock_amd.tune [--out geauxdock_amd/csrc/gd_tiling_table.h]

Then rebuild (python -m geauxdock_amd.build): the library looks a job's
tiling up in that table by size bucket (floor(log2) of ligand atoms L,
protein points P and replicas R; for one-job tilings the ligand's 4-atom
chunk count instead of L's octave), so a tiling is a function of the job
alone.

  shared launches (batches, serving): for each (L, P) bucket of the screening
      sizes (L 12-80, P 1,500-9,000), gd_autotune_shared on a batch of 16
      complexes of a screen (one target, ligands of the bucket, 10 x 11 x 4
      replicas), 32 MC steps per candidate: protein tile points x persistent
      waves per CU.  The waves per CU most buckets pick is kept for all.
  one job per launch: for each (chunks, P, R) bucket of the configurations'
      plain jobs (L 12-80 in 4-atom chunks, R = 25, 220, 440, 880, 1,760),
      gd_autotune on a synthetic complex on the library's default streams, 128
      MC steps per candidate: atoms per lane x tile count, ranked by MC step
      period.

The sweep candidates are the C functions' own (geauxdock_amd/csrc/gd_batch.hip).
"""
import argparse
import collections
import ctypes as C
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import geauxdock_amd as g  # noqa: E402
from geauxdock_amd import abi  # noqa: E402

L_REP = {3: 14, 4: 24, 5: 46, 6: 72}          # representative L of each log2 bucket (12-80)
# one-job tilings are keyed on the ligand's 4-atom chunk count, ceil(L / 4)
L_REPS = {c: (4 * c,) for c in range(3, 21)}     # L = 12 .. 80
P_REP = {10: 1700, 11: 3000, 12: 6000, 13: 9000}  # representative P of each log2 bucket (1,500-9,000)
R_SHAPES = [(5, 5, 1), (5, 11, 4), (10, 11, 4), (20, 11, 4), (20, 11, 8)]   # 25 .. 1,760 replicas


def lg2(x):
    b = 0
    while (2 << b) <= x:
        b += 1
    return b


def chunk_key(L):
    """The library's key of one-job tilings: 4-atom chunks of the ligand."""
    return (max(1, L) + 3) // 4


def tune_shared(ene, steps=32, n_jobs=16):
    entries, waves = [], collections.Counter()
    for lgL, L in L_REP.items():
        for lgP, P in P_REP.items():
            hosts = [g.synth_screen(P, 7000 + lgP, max(12, min(80, L + (k % 5) - 2)), 100 * lgL + k, 10, 11, 4,
                                    ene=ene) for k in range(n_jobs)]
            devs = [g.DeviceJob(h) for h in hosts]
            arr = (C.c_void_p * n_jobs)(*[d.h.value for d in devs])
            t = abi.gd_tiling()
            ms = C.c_double()
            t0 = time.perf_counter()
            g.check(g.lib().gd_autotune_shared(arr, n_jobs, steps, C.byref(t), C.byref(ms)),
                    "gd_autotune_shared")
            for d in devs:
                d.close()
            entries.append(("GD_TUNE_SHARED", lgL, lgP, 0, 4, t.tile_points))
            waves[t.waves_per_cu] += 1
            print("shared L=%d P=%d: %d points, %d waves/CU (%.2f ms; sweep %.1f s)"
                  % (L, P, t.tile_points, t.waves_per_cu, ms.value, time.perf_counter() - t0), flush=True)
    w = waves.most_common(1)[0][0]
    entries.append(("GD_TUNE_WAVES", 0, 0, 0, 4, w))
    return entries


def _period(dev, lc, pt, steps):
    """MC step period (ms) of a device job at one tiling."""
    dev.set_tiling(lc, pt)
    _, st = dev.dock(g.make_mc(steps, move_mode="gaussian", seed=1), stats=True)
    return st.ms_step_period


def tune_plain(ene, steps=128):
    """Each (chunks, P, R) bucket on its representative ligand sizes (one per
    chunk count: L = 4 chunks) at the bucket's P and R: gd_autotune on each;
    where several are given and their winners differ, every winner is timed
    on every job and the one with the smallest summed period relative to each
    job's own best is kept."""
    entries = []
    for hL, Ls in L_REPS.items():
        assert all(chunk_key(L) == hL for L in Ls), Ls
        for lgP, P in P_REP.items():
            for nl, np_, nt in R_SHAPES:
                devs, wins = [], []
                for k, L in enumerate(Ls):
                    h = g.synth_complex(P, L, nl, np_, nt, seed=31 * hL + lgP + 1000 * k, ene=ene)
                    d = g.DeviceJob(h)
                    t = abi.gd_tiling()
                    ms = C.c_double()
                    g.check(g.lib().gd_autotune(d.h, steps, C.byref(t), C.byref(ms)), "gd_autotune")
                    devs.append(d)
                    wins.append(((t.atoms_per_lane, t.tile_points), ms.value))
                cands = sorted({w[0] for w in wins})
                if len(cands) == 1:
                    best = cands[0]
                else:
                    score = {c: sum(_period(d, c[0], c[1], steps) / w[1] for d, w in zip(devs, wins))
                             for c in cands}
                    best = min(cands, key=lambda c: score[c])
                for d in devs:
                    d.close()
                n_tiles = -(-P // best[1])
                R = nl * np_ * nt
                entries.append(("GD_TUNE_PLAIN", hL, lgP, lg2(R), best[0], n_tiles))
                print("plain L=%s P=%d R=%d: %d atoms/lane, %d tiles of %d points (winners %s)"
                      % ("/".join(map(str, Ls)), P, R, best[0], n_tiles, best[1],
                         ", ".join("%d x %d: %.4f ms" % (w[0][0], w[0][1], w[1]) for w in wins)), flush=True)
    return entries


HEADER = """// gd_tiling_table.h -- tiling table of the pair kernel, written by
// `python -m geauxdock_amd.tune` from sweeps on an MI355X (gd_autotune and
// gd_autotune_shared).  A job's tiling is looked up here by size bucket
// (floor(log2) of its ligand atoms L, protein points P and replicas R; for
// one-job tilings the ligand's 4-atom chunk count ceil(L / 4) in place of
// L's), so it is a function of the job alone, never of what else the process
// ran.
//   GD_TUNE_PLAIN  (chunks, lgP, lgR) -> atoms per lane, protein tile count
//   GD_TUNE_SHARED (lgL, lgP)      -> protein tile points (4 atoms per lane)
//   GD_TUNE_WAVES  ()              -> persistent pair kernel waves per CU
// Buckets without an entry use the built-in rule (gd_runtime.hip).
#pragma once

enum { GD_TUNE_PLAIN = 0, GD_TUNE_SHARED = 1, GD_TUNE_WAVES = 2 };

struct GdTuned {
  int mode, lgL, lgP, lgR;
  int lc;     // atoms per lane (GD_TUNE_PLAIN)
  int value;  // tile count (PLAIN), tile points (SHARED), waves per CU (WAVES)
};

// clang-format off
static const GdTuned kTunedTable[] = {
%s
};
// clang-format on
"""


def read_table(path, modes):
    """Entries of the given modes from an existing table header."""
    import re
    out = []
    for m in re.finditer(r"\{(GD_TUNE_\w+), (-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+)\},", open(path).read()):
        if m.group(1) in modes:
            out.append((m.group(1),) + tuple(int(x) for x in m.groups()[1:]))
    return out


def write_table(entries, path, device):
    rows = ["  // %s" % device]
    rows += ["  {%s, %d, %d, %d, %d, %d}," % e for e in entries]
    with open(path, "w") as f:
        f.write(HEADER % "\n".join(rows))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "gd_tiling_table.h"))
    ap.add_argument("--shared-only", action="store_true",
                    help="re-tune the shared-launch entries, keeping the one-job entries of the table the "
                         "library was built with")
    ap.add_argument("--plain-only", action="store_true",
                    help="re-tune the one-job entries, keeping the shared-launch entries of the "
                         "table the library was built with")
    args = ap.parse_args()
    g.check(g.lib().gd_init(0), "gd_init")
    name = C.create_string_buffer(128)
    ncu, clk, hbm = C.c_int(), C.c_int(), C.c_size_t()
    g.lib().gd_device_info(name, 128, C.byref(ncu), C.byref(clk), C.byref(hbm))
    ene = g.load_enepara()
    if args.plain_only and args.shared_only:
        ap.error("--plain-only and --shared-only exclude each other")
    src = os.path.join(HERE, "csrc", "gd_tiling_table.h")   # the table the library was built with
    if args.plain_only:
        entries = read_table(src, ("GD_TUNE_SHARED", "GD_TUNE_WAVES"))
        print("kept %d shared-launch entries of %s" % (len(entries), src), flush=True)
    else:
        entries = tune_shared(ene)
    if args.shared_only:
        kept = read_table(src, ("GD_TUNE_PLAIN",))
        print("kept %d one-job entries of %s" % (len(kept), src), flush=True)
        entries += kept
    else:
        entries += tune_plain(ene)
    write_table(entries, args.out, "tuned on %s" % name.value.decode())
    print("wrote %s (%d entries)" % (args.out, len(entries)), flush=True)


if __name__ == "__main__":
    main()
