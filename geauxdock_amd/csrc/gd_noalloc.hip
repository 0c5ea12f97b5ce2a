// gd_noalloc.hip -- gd_energy_eval_noalloc: the energy of arbitrary poses with
// no allocation at all, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// it).  The kernels' scratch -- the evaluation's copy of the job struct, its
// work items, the staged poses, transformed ligands, partial sums and results
// -- lives in one file-scope device array of GD_NOALLOC_SCRATCH_BYTES, held by
// every GPU the library runs on until the process ends.
//
// Same kernels and the same tiling as gd_energy_eval, so the results are the
// same bits.  Poses run in chunks of whole 64-lane blocks that fit the
// scratch; the chunk's work items are gathered from the job's own list (item
// order: chunk, tile group, replica block innermost), so no host-side list is
// built either.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU race (the array keeps the last call's contents).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/geauxdock_amd.h"
#include "gd_device.h"
#include "gd_internal.h"

#define GD_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[GD_NOALLOC_SCRATCH_BYTES];

namespace {

// The device address of `scratch` on each device, from hipGetSymbolAddress
// with that device current (each GPU holds its own copy of the array).
void *g_scratch[GD_NOALLOC_MAX_DEVICES];
std::mutex g_scratch_mu;

gd_status scratch_ptr(void **out) {
  const int dev = gdr::device();
  if (dev < 0 || dev >= GD_NOALLOC_MAX_DEVICES) return GD_ERR_STATE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return gdr::hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return GD_OK;
}

inline size_t up256(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of scratch one evaluation chunk of nb 64-lane blocks takes.
struct Carve {
  size_t job, items, conf, pose, ligi, prti, prop, coords, plin, phpc, pmcs, fin, evt, evr, total;
};
Carve carve(const GdJob &g, int nb, int n_items) {
  const size_t Rp = (size_t)nb * 64;
  Carve c;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += up256(bytes);
    return at;
  };
  c.job = take(sizeof(GdJob));
  c.items = take((size_t)n_items * sizeof(GdItem));
  c.conf = take(Rp * 2 * sizeof(int));
  c.pose = take(Rp * 6 * sizeof(float));
  c.ligi = take(Rp * sizeof(int));
  c.prti = take(Rp * sizeof(int));
  c.prop = take(6 * Rp * sizeof(float));
  c.coords = take((size_t)3 * g.L * Rp * sizeof(float));
  c.plin = take((size_t)GD_NLIN * g.nL * g.nG * Rp * sizeof(float));
  c.phpc = take((size_t)g.L * g.nG * Rp * sizeof(float));
  c.pmcs = take((size_t)(g.M > 0 ? g.M : 1) * g.nL * Rp * sizeof(float));
  c.fin = take((size_t)nb * GD_RK * 8 * 64 * sizeof(float));
  c.evt = take(10 * Rp * sizeof(float));
  c.evr = take(9 * Rp * sizeof(float));
  c.total = o;
  return c;
}

}  // namespace

// Lanes of an evaluation chunk: conformers and poses of the first cnt lanes
// from the staged inputs (pose [cnt][6] -> prop [6][Rp]), the padding lanes
// conformer 0 at the zero pose; and the chunk's work items, gathered from the
// job's list of nb_job blocks per (chunk, tile group).
__global__ __launch_bounds__(256) void gd_noalloc_setup_kernel(
    const int *__restrict__ conf, const float *__restrict__ pose, int cnt, int Rp,
    int *__restrict__ ligi, int *__restrict__ prti, float *__restrict__ prop,
    const GdItem *__restrict__ job_items, int nb_job, GdItem *__restrict__ items, int nb,
    int n_items) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < Rp) {
    const bool live = i < cnt;
    ligi[i] = live ? conf[2 * i] : 0;
    prti[i] = live ? conf[2 * i + 1] : 0;
    for (int k = 0; k < 6; ++k) prop[(size_t)k * Rp + i] = live ? pose[6 * (size_t)i + k] : 0.0f;
  }
  if (i < n_items) items[i] = job_items[(size_t)(i / nb) * nb_job + (i % nb)];
}

extern "C" gd_status gd_energy_eval_noalloc(gd_job *j, int n, const int *conf, const float *pose,
                                            float *terms, float *raw) {
  if (!j || n < 0 || (n > 0 && (!conf || !pose || !terms))) return GD_ERR_ARG;
  if (n == 0) return GD_OK;
  gd_status st = gdr::ensure_device();
  if (st != GD_OK) return st;
  const GdJob &g = j->h;
  for (int i = 0; i < n; ++i)
    if (conf[2 * i] < 0 || conf[2 * i] >= g.n_lig || conf[2 * i + 1] < 0 || conf[2 * i + 1] >= g.n_prt)
      return GD_ERR_ARG;
  void *base = nullptr;
  if ((st = scratch_ptr(&base)) != GD_OK) return st;
  const int nb_job = g.Rp / 64;
  const int per_block_items = g.nL * g.nG;  // items per replica block
  // the largest chunk (whole blocks, at most the job's) that fits the scratch
  int nb = nb_job;
  while (nb > 0 && carve(g, nb, per_block_items * nb).total > (size_t)GD_NOALLOC_SCRATCH_BYTES) --nb;
  if (nb == 0) {
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "gd_energy_eval_noalloc: one 64-pose block of this job needs %zu bytes of scratch, "
                  "more than the %d-byte static workspace",
                  carve(g, 1, per_block_items).total, (int)GD_NOALLOC_SCRATCH_BYTES);
    gdr::set_error(buf);
    return GD_ERR_SIZE;
  }
  hipStream_t s = j->stream;
  hipError_t e = hipStreamSynchronize(s);  // the job's own work first
  if (e != hipSuccess) return gdr::hip_fail(e, "hipStreamSynchronize");
  char *sb = (char *)base;
  for (int at = 0; at < n; at += nb * 64) {
    const int cnt = n - at < nb * 64 ? n - at : nb * 64;
    const int nbc = (cnt + 63) / 64;
    const int Rp = nbc * 64, n_items = per_block_items * nbc;
    const Carve c = carve(g, nbc, n_items);
    // the evaluation's job struct: the job's tables, lanes and buffers in scratch
    GdJob x = g;
    x.R = cnt;
    x.Rp = Rp;
    x.ligi = (int *)(sb + c.ligi);
    x.prti = (int *)(sb + c.prti);
    x.prop = (float *)(sb + c.prop);
    x.coords = (float *)(sb + c.coords);
    x.part_lin = (float *)(sb + c.plin);
    x.part_hpc = (float *)(sb + c.phpc);
    x.part_mcs = (float *)(sb + c.pmcs);
    x.fin_part = (float *)(sb + c.fin);
    x.ev_terms = (float *)(sb + c.evt);
    x.ev_raw = (float *)(sb + c.evr);
    x.rec = nullptr;
    x.rec_cap = 0;
    x.record = 0;
    const GdJob *d_x = (const GdJob *)(sb + c.job);
    GdItem *d_it = (GdItem *)(sb + c.items);
    if ((e = hipMemcpyAsync((void *)d_x, &x, sizeof x, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(sb + c.conf, conf + 2 * (size_t)at, (size_t)cnt * 2 * sizeof(int),
                            hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(sb + c.pose, pose + 6 * (size_t)at, (size_t)cnt * 6 * sizeof(float),
                            hipMemcpyHostToDevice, s)) != hipSuccess)
      return gdr::hip_fail(e, "gd_energy_eval_noalloc staging");
    const int nthr = Rp > n_items ? Rp : n_items;
    hipLaunchKernelGGL(gd_noalloc_setup_kernel, dim3((nthr + 255) / 256), dim3(256), 0, s,
                       (const int *)(sb + c.conf), (const float *)(sb + c.pose), cnt, Rp, x.ligi,
                       x.prti, x.prop, (const GdItem *)j->d_items, nb_job, d_it, nbc, n_items);
    if ((e = hipGetLastError()) != hipSuccess) return gdr::hip_fail(e, "gd_noalloc_setup_kernel");
    // the job's block list starts with blocks 0, 64, ...: its first nbc entries
    if ((st = gdr::launch_transform(d_x, j->d_blocks, nbc, s)) != GD_OK ||
        (st = gdr::launch_pair(g.lc, d_x, d_it, n_items, s)) != GD_OK ||
        (st = gdr::launch_finish(d_x, j->d_blocks, nbc, GD_MODE_EVAL, 0, s)) != GD_OK)
      return st;
    if ((e = hipMemcpyAsync(terms + 10 * (size_t)at, x.ev_terms, (size_t)cnt * 10 * sizeof(float),
                            hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (raw && (e = hipMemcpyAsync(raw + 9 * (size_t)at, x.ev_raw, (size_t)cnt * 9 * sizeof(float),
                                    hipMemcpyDeviceToHost, s)) != hipSuccess) ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return gdr::hip_fail(e, "gd_energy_eval_noalloc readback");
  }
  return GD_OK;
}
