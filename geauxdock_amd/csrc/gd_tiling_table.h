// gd_tiling_table.h -- tiling table of the pair kernel, written by
// `python -m geauxdock_amd.tune` from sweeps on an MI355X (gd_autotune and
// gd_autotune_shared).  A job's tiling is looked up here by size bucket
// (floor(log2) of its ligand atoms L, protein points P and replicas R), so it
// is a function of the job alone, never of what else the process ran.
//   GD_TUNE_PLAIN  (lgL, lgP, lgR) -> atoms per lane, protein tile count
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
  // tuned on gfx950:sramecc+:xnack-, 256 CUs
  {GD_TUNE_SHARED, 3, 10, 0, 4, 64},
  {GD_TUNE_SHARED, 3, 11, 0, 4, 96},
  {GD_TUNE_SHARED, 3, 12, 0, 4, 128},
  {GD_TUNE_SHARED, 3, 13, 0, 4, 128},
  {GD_TUNE_SHARED, 4, 10, 0, 4, 64},
  {GD_TUNE_SHARED, 4, 11, 0, 4, 80},
  {GD_TUNE_SHARED, 4, 12, 0, 4, 128},
  {GD_TUNE_SHARED, 4, 13, 0, 4, 128},
  {GD_TUNE_SHARED, 5, 10, 0, 4, 48},
  {GD_TUNE_SHARED, 5, 11, 0, 4, 64},
  {GD_TUNE_SHARED, 5, 12, 0, 4, 128},
  {GD_TUNE_SHARED, 5, 13, 0, 4, 96},
  {GD_TUNE_SHARED, 6, 10, 0, 4, 64},
  {GD_TUNE_SHARED, 6, 11, 0, 4, 96},
  {GD_TUNE_SHARED, 6, 12, 0, 4, 128},
  {GD_TUNE_SHARED, 6, 13, 0, 4, 128},
  {GD_TUNE_WAVES, 0, 0, 0, 4, 32},
  {GD_TUNE_PLAIN, 3, 10, 4, 4, 155},
  {GD_TUNE_PLAIN, 3, 10, 7, 4, 122},
  {GD_TUNE_PLAIN, 3, 10, 8, 4, 122},
  {GD_TUNE_PLAIN, 3, 10, 9, 4, 155},
  {GD_TUNE_PLAIN, 3, 10, 10, 4, 85},
  {GD_TUNE_PLAIN, 3, 11, 4, 4, 188},
  {GD_TUNE_PLAIN, 3, 11, 7, 4, 125},
  {GD_TUNE_PLAIN, 3, 11, 8, 4, 188},
  {GD_TUNE_PLAIN, 3, 11, 9, 4, 158},
  {GD_TUNE_PLAIN, 3, 11, 10, 4, 125},
  {GD_TUNE_PLAIN, 3, 12, 4, 4, 375},
  {GD_TUNE_PLAIN, 3, 12, 7, 4, 375},
  {GD_TUNE_PLAIN, 3, 12, 8, 4, 375},
  {GD_TUNE_PLAIN, 3, 12, 9, 4, 375},
  {GD_TUNE_PLAIN, 3, 12, 10, 4, 128},
  {GD_TUNE_PLAIN, 3, 13, 4, 4, 563},
  {GD_TUNE_PLAIN, 3, 13, 7, 4, 563},
  {GD_TUNE_PLAIN, 3, 13, 8, 4, 563},
  {GD_TUNE_PLAIN, 3, 13, 9, 4, 392},
  {GD_TUNE_PLAIN, 3, 13, 10, 4, 127},
  {GD_TUNE_PLAIN, 4, 10, 4, 4, 155},
  {GD_TUNE_PLAIN, 4, 10, 7, 4, 155},
  {GD_TUNE_PLAIN, 4, 10, 8, 4, 95},
  {GD_TUNE_PLAIN, 4, 10, 9, 4, 95},
  {GD_TUNE_PLAIN, 4, 10, 10, 4, 95},
  {GD_TUNE_PLAIN, 4, 11, 4, 4, 188},
  {GD_TUNE_PLAIN, 4, 11, 7, 4, 158},
  {GD_TUNE_PLAIN, 4, 11, 8, 4, 188},
  {GD_TUNE_PLAIN, 4, 11, 9, 4, 94},
  {GD_TUNE_PLAIN, 4, 11, 10, 4, 94},
  {GD_TUNE_PLAIN, 4, 12, 4, 4, 375},
  {GD_TUNE_PLAIN, 4, 12, 7, 4, 375},
  {GD_TUNE_PLAIN, 4, 12, 8, 4, 375},
  {GD_TUNE_PLAIN, 4, 12, 9, 4, 261},
  {GD_TUNE_PLAIN, 4, 12, 10, 4, 96},
  {GD_TUNE_PLAIN, 4, 13, 4, 4, 563},
  {GD_TUNE_PLAIN, 4, 13, 7, 4, 563},
  {GD_TUNE_PLAIN, 4, 13, 8, 4, 530},
  {GD_TUNE_PLAIN, 4, 13, 9, 4, 158},
  {GD_TUNE_PLAIN, 4, 13, 10, 4, 96},
  {GD_TUNE_PLAIN, 5, 10, 4, 4, 155},
  {GD_TUNE_PLAIN, 5, 10, 7, 4, 155},
  {GD_TUNE_PLAIN, 5, 10, 8, 4, 155},
  {GD_TUNE_PLAIN, 5, 10, 9, 4, 95},
  {GD_TUNE_PLAIN, 5, 10, 10, 4, 71},
  {GD_TUNE_PLAIN, 5, 11, 4, 4, 158},
  {GD_TUNE_PLAIN, 5, 11, 7, 4, 158},
  {GD_TUNE_PLAIN, 5, 11, 8, 4, 188},
  {GD_TUNE_PLAIN, 5, 11, 9, 4, 94},
  {GD_TUNE_PLAIN, 5, 11, 10, 4, 72},
  {GD_TUNE_PLAIN, 5, 12, 4, 4, 375},
  {GD_TUNE_PLAIN, 5, 12, 7, 4, 128},
  {GD_TUNE_PLAIN, 5, 12, 8, 4, 261},
  {GD_TUNE_PLAIN, 5, 12, 9, 4, 96},
  {GD_TUNE_PLAIN, 5, 12, 10, 4, 128},
  {GD_TUNE_PLAIN, 5, 13, 4, 4, 563},
  {GD_TUNE_PLAIN, 5, 13, 7, 4, 450},
  {GD_TUNE_PLAIN, 5, 13, 8, 4, 265},
  {GD_TUNE_PLAIN, 5, 13, 9, 4, 96},
  {GD_TUNE_PLAIN, 5, 13, 10, 4, 127},
  {GD_TUNE_PLAIN, 6, 10, 4, 4, 155},
  {GD_TUNE_PLAIN, 6, 10, 7, 4, 107},
  {GD_TUNE_PLAIN, 6, 10, 8, 4, 95},
  {GD_TUNE_PLAIN, 6, 10, 9, 4, 95},
  {GD_TUNE_PLAIN, 6, 10, 10, 4, 55},
  {GD_TUNE_PLAIN, 6, 11, 4, 4, 188},
  {GD_TUNE_PLAIN, 6, 11, 7, 4, 112},
  {GD_TUNE_PLAIN, 6, 11, 8, 4, 125},
  {GD_TUNE_PLAIN, 6, 11, 9, 4, 72},
  {GD_TUNE_PLAIN, 6, 11, 10, 4, 56},
  {GD_TUNE_PLAIN, 6, 12, 4, 4, 375},
  {GD_TUNE_PLAIN, 6, 12, 7, 4, 112},
  {GD_TUNE_PLAIN, 6, 12, 8, 4, 158},
  {GD_TUNE_PLAIN, 6, 12, 9, 4, 112},
  {GD_TUNE_PLAIN, 6, 12, 10, 4, 87},
  {GD_TUNE_PLAIN, 6, 13, 4, 4, 563},
  {GD_TUNE_PLAIN, 6, 13, 7, 4, 311},
  {GD_TUNE_PLAIN, 6, 13, 8, 4, 158},
  {GD_TUNE_PLAIN, 6, 13, 9, 4, 127},
  {GD_TUNE_PLAIN, 6, 13, 10, 4, 88},
};
// clang-format on
