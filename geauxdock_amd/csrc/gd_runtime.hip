// gd_runtime.hip -- host runtime of the C ABI: device selection, packing a
// complex into the device layout of gd_device.h, tiling, kernel launches,
// the plain one-job docking path, energy evaluation and accept records.
// Batching and serving build on the same job objects (gd_batch.hip).
//
// Reference call sequence replaced (src/backend_gpu/backend_gpu.cu:64-133,
// prepair.C:12-135): CopyH2D of a 6.3 MiB Complex per job, one launch for
// the initial energies, one launch + blocking D2H of records per dump.  Here
// a job's tables are packed once into device memory sized to the job, the
// replica state never leaves the device, and every MC step is a pair-,
// reduce- and finish-kernel launch on the job's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/geauxdock_amd.h"
#include "gd_device.h"
#include "gd_internal.h"
#include "gd_tiling_table.h"

// kernels (gd_kernels.hip)
template <int LC>
__global__ void gd_pair_kernel(const GdJob *, const GdItem *, int);
__global__ void gd_transform_kernel(const GdJob *, const GdBlock *, int *);
__global__ void gd_reduce_kernel(const GdJob *, const GdBlock *);
__global__ void gd_finish_kernel(const GdJob *, const GdBlock *, int, int, const int *, int *);
__global__ void gd_exchange_kernel(const GdJob *, const GdExch *);
__global__ void gd_move_draws_kernel(uint64_t, uint32_t, int, GdScale6, int, uint32_t, int, float *);
__global__ void gd_clock_kernel(long long, float, long long *, float *);
__global__ void gd_copy_kernel(const float4 *, float4 *, size_t);
__global__ void gd_swap_propose_kernel(const GdJob *, const GdBlock *, int, const int *);
__global__ void gd_swap_accept_kernel(const GdJob *, const GdBlock *, int, const int *, int,
                                      const int *);

namespace gdr {

static thread_local std::string g_err;
static int g_device = -1;
static int g_cu = 0;
static std::mutex g_mu;

void set_error(const std::string &s) { g_err = s; }
const char *last_error() { return g_err.c_str(); }

gd_status hip_fail(hipError_t e, const char *what) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  set_error(buf);
  return GD_ERR_HIP;
}

KTimer::~KTimer() {
  for (auto e : ev) hipEventDestroy(e);
}
gd_status KTimer::begin(hipStream_t s) {
  if (used + 2 > (int)ev.size()) {
    hipEvent_t a, b;
    hipError_t e;
    // timing-only events: without the system-scope fence a record costs no
    // L2 writeback and invalidate (about 5 us per record on MI355X, paid by
    // the work around it); the elapsed times are read after a stream
    // synchronisation, so no event needs to order memory
    if ((e = hipEventCreateWithFlags(&a, hipEventDisableSystemFence)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b, hipEventDisableSystemFence)) != hipSuccess)
      return hip_fail(e, "hipEventCreateWithFlags");
    ev.push_back(a);
    ev.push_back(b);
  }
  hipError_t e = hipEventRecord(ev[used], s);
  return e == hipSuccess ? GD_OK : hip_fail(e, "hipEventRecord");
}
gd_status KTimer::end(hipStream_t s) {
  hipError_t e = hipEventRecord(ev[used + 1], s);
  used += 2;
  return e == hipSuccess ? GD_OK : hip_fail(e, "hipEventRecord");
}
gd_status KTimer::collect() {
  for (int i = 0; i + 1 < used; i += 2) {
    float t = 0.0f;
    hipError_t e = hipEventElapsedTime(&t, ev[i], ev[i + 1]);
    if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
    ms += t;
    ++n;
  }
  if (used >= 4) {
    float t = 0.0f;
    hipError_t e = hipEventElapsedTime(&t, ev[0], ev[used - 2]);
    if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
    first_to_last_ms += t;
  }
  used = 0;
  return GD_OK;
}

// Pair-phase timer of the multi-stream plain path: per sampled MC step, the
// span from the first start to the last end of its parts' launches (events on
// every part's stream), every event timed from one reference recorded earlier
// on the first stream (so no elapsed time is negative).
struct SpanTimer {
  std::vector<hipEvent_t> ev;
  int used = 0;
  hipEvent_t ref = nullptr;
  double ms = 0.0;
  int n = 0;
  ~SpanTimer() {
    for (auto e : ev) hipEventDestroy(e);
    if (ref) hipEventDestroy(ref);
  }
  hipEvent_t next() {
    if (used == (int)ev.size()) {
      hipEvent_t e;
      if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
      ev.push_back(e);
    }
    return ev[used++];
  }
  gd_status mark_ref(hipStream_t s) {
    if (!ref && hipEventCreateWithFlags(&ref, hipEventDisableSystemFence) != hipSuccess)
      return hip_fail(hipErrorOutOfMemory, "hipEventCreate");
    const hipError_t e = hipEventRecord(ref, s);
    return e == hipSuccess ? GD_OK : hip_fail(e, "hipEventRecord");
  }
  gd_status record(hipStream_t s) {  // per sample: start and end of every part's launch
    hipEvent_t e = next();
    if (!e) return hip_fail(hipErrorOutOfMemory, "hipEventCreate");
    const hipError_t r = hipEventRecord(e, s);
    return r == hipSuccess ? GD_OK : hip_fail(r, "hipEventRecord");
  }
  double parts_ms = 0.0, first_a = -1.0, last_a = -1.0;
  int n_part_launches = 0;
  gd_status collect(int np) {  // after every part's stream is synchronised
    for (int i = 0; i + 2 * np - 1 < used; i += 2 * np) {
      float t0 = 0.0f, t_lo = 0.0f, t_hi = 0.0f;
      for (int k = 0; k < np; ++k) {
        float a, b;
        hipError_t e = hipEventElapsedTime(&a, ref, ev[i + 2 * k]);
        if (e == hipSuccess) e = hipEventElapsedTime(&b, ref, ev[i + 2 * k + 1]);
        if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
        if (k == 0) t0 = t_lo = a, t_hi = b;
        t_lo = std::min(t_lo, a);
        t_hi = std::max(t_hi, b);
        parts_ms += b - a;
        ++n_part_launches;
      }
      ms += t_hi - t_lo;
      ++n;
      if (first_a < 0.0) first_a = t0;
      last_a = t0;
    }
    used = 0;
    return GD_OK;
  }
};

int device() { return g_device; }
int n_cu() { return g_cu; }

gd_status ensure_device() {
  if (g_device >= 0) {
    hipError_t e = hipSetDevice(g_device);
    return e == hipSuccess ? GD_OK : hip_fail(e, "hipSetDevice");
  }
  return gd_init(0);
}

// ---- device arena: one hipMalloc per job, carved into 256-byte aligned pieces
struct Arena {
  size_t off = 0;
  std::vector<std::pair<void **, size_t>> reqs;
  template <class T>
  void add(T **p, size_t n) {
    reqs.push_back({(void **)p, (n * sizeof(T) + 255) / 256 * 256});
    off += reqs.back().second;
  }
  gd_status commit(void **base) {
    *base = nullptr;
    if (off == 0) return GD_OK;
    hipError_t e = hipMalloc(base, off);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    size_t o = 0;
    for (auto &r : reqs) {
      *r.first = (char *)*base + o;
      o += r.second;
    }
    return GD_OK;
  }
};

// Exact float threshold T on r^2 such that sqrtf(r2) <= pmf0  <=>  r2 <= T
// (the reference tests dst - pmf0 <= 0 with dst = sqrtf(r2),
// energy_prt_v1.C:88,100).  sqrtf is correctly rounded and monotone.
static float sqrt_threshold(float pmf0) {
  if (!(pmf0 >= 0.0f)) return -1.0f;  // never true for r2 >= 0
  float x = (float)((double)pmf0 * (double)pmf0);
  while (std::sqrt(x) > pmf0) x = std::nextafter(x, 0.0f);
  for (;;) {
    const float n = std::nextafter(x, INFINITY);
    if (std::sqrt(n) <= pmf0 && std::isfinite(n))
      x = n;
    else
      break;
  }
  return x;
}

// Far-field threshold on r^2 for one (ligand type, point type) pair: from it on
// the PMF exponent (6 - r/2)(r - pmf0) log2(e) stays below -26 (so 1 + exp2 of it
// rounds to 1 in FP32 even with the kernel's rounding of the exponent) and
// el1 r >= 1 (the electrostatic factor is 1/(el1 r)); both hold for every larger r
// because the exponent decreases beyond r = 6 + pmf0/2.  Rounded up.
float far_r2_threshold(double pmf0, double el1) {
  if (!(el1 > 0.0)) return std::numeric_limits<float>::infinity();
  const double log2e = 1.44269504088896340736;
  // -log2(e)/2 r^2 + log2(e)(6 + pmf0/2) r - 6 log2(e) pmf0 + 26 = 0, larger root
  const double qa = -0.5 * log2e, qb = log2e * (6.0 + 0.5 * pmf0), qc = -6.0 * log2e * pmf0 + 26.0;
  const double disc = qb * qb - 4.0 * qa * qc;
  double r = disc > 0.0 ? (-qb - std::sqrt(disc)) / (2.0 * qa) : 6.0 + 0.5 * pmf0;
  r = std::max(r, 6.0 + 0.5 * pmf0) * (1.0 + 1e-6);
  while (log2e * (6.0 - 0.5 * r) * (r - pmf0) > -26.0) r *= 1.0 + 1e-6;
  r = std::max(r, (1.0 / el1) * (1.0 + 1e-6));
  const float f = (float)(r * r);
  return (double)f >= r * r ? f : std::nextafter(f, std::numeric_limits<float>::infinity());
}

// The row threshold of the pair loop's near branch: the far-field threshold
// raised to cover the terms that branch also holds, so that a wave whose lanes
// are all beyond it skips exact zeros only:
//   hydrophobic, r^2 <= 81; psp, r^2 <= psp_thr (negative: no psp term);
//   H-bond (hb), hdb1 2^-q with q = ((r - hdb0) c)^2, c = hdb1 sqrt(log2(e)/2):
//   2^-q is exactly 0 in FP32 for q >= 151, and q >= 160 from r = hdb0 + sqrt(160) / c on.
static float near_r2_threshold(double pmf0, double el1, float psp_thr, bool hb, double hdb0,
                               double c) {
  double r2n = far_r2_threshold(pmf0, el1);
  r2n = std::max(r2n, 81.0 * (1.0 + 1e-6));
  if (psp_thr >= 0.0f) r2n = std::max(r2n, (double)psp_thr * (1.0 + 1e-6));
  if (hb && c > 0.0) {
    const double rh = hdb0 + std::sqrt(160.0) / c;
    r2n = std::max(r2n, rh * rh * (1.0 + 1e-5));
  }
  float f = (float)r2n;
  if ((double)f < r2n) f = std::nextafter(f, std::numeric_limits<float>::infinity());
  return f;
}

// ---------------------------------------------------------------------------
// Tiling: ligand chunk size lc (kernel template), protein tile ptile.  Looked
// up by size bucket in the tuned table (gd_tiling_table.h), else the rule
// below; either way a function of the job (and the device's CU count) only, so
// a job's reduction order, and hence its bits, is the same in every process,
// alone or in a batch.
int lg2_floor(int x) {
  int b = 0;
  while (b < 30 && (2 << b) <= x) ++b;
  return b;
}

// One-job tilings are keyed on the ligand's 4-atom chunk count, ceil(L / 4):
// on two streams the best tile count keeps each half-launch's items (chunks x
// tile groups x replica blocks) near the machine's group slots, so it follows
// the chunk count exactly rather than a size class of L.
int chunk_key(int L) { return (std::max(1, L) + 3) / 4; }

static const GdTuned *tuned_entry(int mode, int L, int P, int R) {
  const int lgL = mode == GD_TUNE_PLAIN ? chunk_key(L) : lg2_floor(std::max(1, L)),
            lgP = lg2_floor(std::max(1, P)), lgR = lg2_floor(std::max(1, R));
  for (const GdTuned &t : kTunedTable) {
    if (t.mode != mode) continue;
    if (mode == GD_TUNE_WAVES ||
        (t.lgL == lgL && t.lgP == lgP && (mode == GD_TUNE_SHARED || t.lgR == lgR)))
      return &t;
  }
  return nullptr;
}

// ptile giving at least nP tiles of P points (the kernel rounds nP up to a
// multiple of GD_GROUP)
static int ptile_for_tiles(int P, int nP) {
  nP = std::max(1, std::min(nP, P));
  return (P + nP - 1) / nP;
}

void default_tiling(int L, int P, int R, int &lc, int &ptile) {
  if (const GdTuned *t = tuned_entry(GD_TUNE_PLAIN, L, P, R)) {
    lc = t->lc;
    ptile = ptile_for_tiles(P, t->value);
    return;
  }
  // 4 atoms per wave keeps the pair kernel at ~62 VGPRs (8 waves per SIMD);
  // ~22 waves per SIMD over the launch (about three waves per slot) balanced
  // per-tile overhead (KDE and MCS shares, partial sums) against the launch's
  // tail in a sweep of tile counts on a 3000-point, 1760-replica job
  lc = 4;
  const int nL = (L + lc - 1) / lc;
  const int nblk = (R + 63) / 64;
  const int simds = 4 * (g_cu > 0 ? g_cu : 256);
  const int target_waves = 22 * simds;
  int nP = std::max(1, target_waves / std::max(1, nblk * nL));
  nP = std::min(nP, std::max(1, P / 16));
  nP = (nP + GD_GROUP - 1) / GD_GROUP * GD_GROUP;
  ptile = (P + nP - 1) / nP;
}

// Tiling of a job inside shared launches (batches, serving): 4 atoms per lane
// (one kernel template for every job of a launch) and the bucket's tile size,
// 64 points where the table has none.
int shared_ptile(int L, int P) {
  if (const GdTuned *t = tuned_entry(GD_TUNE_SHARED, L, P, 0)) return std::min(P, t->value);
  return std::min(P, 64);
}

// Resident waves per CU of the persistent pair kernel for lc atoms per lane:
// the tuned value for 4 atoms, else as many 8-wave groups as the registers
// allow (4 atoms: ~60 VGPRs, 8 waves per SIMD).
int pair_waves_per_cu(int lc) {
  if (lc == 4)
    if (const GdTuned *t = tuned_entry(GD_TUNE_WAVES, 0, 0, 0)) return t->value;
  return lc == 4 ? 32 : lc == 8 ? 20 : 8;
}

static gd_status alloc_partials(gd_job *j, GdJob &g, int cap_rp) {
  // partials depend on the tiling; they live in their own allocation
  if (j->d_part) {
    hipFree(j->d_part);
    j->d_part = nullptr;
  }
  Arena a;
  const size_t nS = (size_t)g.nL * g.nG;
  a.add(&g.part_lin, (size_t)GD_NLIN * nS * cap_rp);
  a.add(&g.part_hpc, (size_t)g.L * g.nG * cap_rp);
  a.add(&g.part_mcs, (size_t)std::max(1, g.M) * g.nL * cap_rp);
  a.add(&g.fin_part, (size_t)(cap_rp / 64) * GD_RK * 8 * 64);
  gd_status st = a.commit(&j->d_part);
  j->part_bytes = st == GD_OK ? a.off : 0;
  return st;
}

// Device order of the protein points for the job's tile count: class by
// class, tile b of nP reads points pcls[c] + b, + nP, ... (strided, see the
// pair kernel).  The class's points sorted by (type, spread) are dealt out so
// that each tile's points are one contiguous piece of that list: few type
// changes per tile (the pair kernel keeps a chunk's force-field rows across a
// type), and, by the spread order, a representative mix of distances.  The
// order changes only summation order.
static gd_status layout_points(gd_job *j) {
  GdJob &g = j->h;
  const int P = g.P, nP = g.nP, np = g.n_prt;
  std::vector<int> ord(P), porder(P);
  for (int p = 0; p < P; ++p) ord[p] = p;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
    return j->src_key[x] != j->src_key[y] ? j->src_key[x] < j->src_key[y]
                                          : j->src_spread[x] < j->src_spread[y];
  });
  for (int c = 0; c < 4; ++c) {
    const int lo = g.pcls[c], n = g.pcls[c + 1] - lo, q = n / nP, r = n % nP;
    for (int i = 0; i < n; ++i) {
      int b, k;  // column (tile) and row of list element i: columns are filled in turn
      if (i < r * (q + 1)) {
        b = i / (q + 1), k = i % (q + 1);
      } else {
        const int i2 = i - r * (q + 1);
        b = r + i2 / q, k = i2 % q;
      }
      porder[lo + k * nP + b] = ord[lo + i];
    }
  }
  const size_t nty1 = j->src_pspv.size() / (size_t)P;
  std::vector<float4> pco((size_t)np * P), pattr(P);
  std::vector<float> pspv(j->src_pspv.size());
  for (int i = 0; i < np; ++i)
    for (int q2 = 0; q2 < P; ++q2) pco[(size_t)i * P + q2] = j->src_pco[(size_t)i * P + porder[q2]];
  for (int q2 = 0; q2 < P; ++q2) pattr[q2] = j->src_pattr[porder[q2]];
  for (size_t t = 0; t < nty1; ++t)
    for (int q2 = 0; q2 < P; ++q2) pspv[t * P + q2] = j->src_pspv[t * P + porder[q2]];
  hipError_t e;
  if ((e = hipStreamSynchronize(j->stream)) != hipSuccess ||
      (e = hipMemcpyAsync((void *)g.pcoord, pco.data(), pco.size() * sizeof(float4),
                          hipMemcpyHostToDevice, j->stream)) != hipSuccess ||
      (e = hipMemcpyAsync((void *)g.pattr, pattr.data(), pattr.size() * sizeof(float4),
                          hipMemcpyHostToDevice, j->stream)) != hipSuccess ||
      (e = hipMemcpyAsync((void *)g.pspv, pspv.data(), pspv.size() * sizeof(float),
                          hipMemcpyHostToDevice, j->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(j->stream)) != hipSuccess)
    return hip_fail(e, "point layout upload");
  j->laid_nP = nP;
  return GD_OK;
}

gd_status apply_tiling(gd_job *j, int lc, int ptile) {
  GdJob &g = j->h;
  if (lc != 4 && lc != 8 && lc != 16) return GD_ERR_ARG;
  ptile = std::max(1, std::min(ptile, g.P));
  g.lc = lc;
  g.nL = (g.L + lc - 1) / lc;
  g.ptile = ptile;
  // protein tiles (tile b takes the points b, b + nP, ... of each class): a
  // multiple of GD_GROUP so that a work group's waves take tiles 8g .. 8g + 7
  const int nP = (g.P + ptile - 1) / ptile;
  g.nP = (nP + GD_GROUP - 1) / GD_GROUP * GD_GROUP;
  g.nG = g.nP / GD_GROUP;
  gd_status st = alloc_partials(j, g, g.Rp);
  if (st != GD_OK) return st;
  if (g.nP != j->laid_nP && (st = layout_points(j)) != GD_OK) return st;
  // One item per work group: (replica block, chunk, tile group).  Chunks go in
  // order of decreasing KDE work (anchors of their atoms' bins), the only part
  // of an item's cost that differs between chunks, so the launch ends with its
  // cheapest items; within a chunk, tile groups, then replica blocks, so items
  // dispatched together read the same force-field rows and KDE anchors.  The
  // order changes no result: every item writes its own partial-sum slot.
  // (Measured on a 3,000-point, 1,760-replica job: 117.3 us against 119.8 us for
  // replica-block-major order, and 121.0 us for the reverse chunk order.)
  std::vector<int> chunk(g.nL), wkde(g.nL, 0);
  for (int a = 0; a < g.nL; ++a) {
    chunk[a] = a;
    for (int l = a * lc; l < std::min((a + 1) * lc, g.L); ++l) wkde[a] += j->atom_kde[l];
  }
  std::stable_sort(chunk.begin(), chunk.end(), [&](int x, int y) { return wkde[x] > wkde[y]; });
  j->items.clear();
  for (int a : chunk)
    for (int gi = 0; gi < g.nG; ++gi)
      for (int rb = 0; rb < g.Rp; rb += 64) j->items.push_back({0, rb, a, gi});
  return GD_OK;
}

// ---------------------------------------------------------------------------
// Pack a complex into the device layout.
static gd_status pack_job(const gd_job_desc *d, gd_job *j) {
  const int L = d->lig_natom, P = d->prt_npoint, K = d->kde_npoint, M = d->mcs_nrow;
  const int nl = d->n_lig, np = d->n_prt, nt = d->n_tmp;
  if (L < 1 || P < 1 || K < 0 || M < 0 || nl < 1 || np < 1 || nt < 1) return GD_ERR_ARG;
  if (!d->lig_x || !d->lig_y || !d->lig_z || !d->lig_t || !d->lig_c || !d->prt_x || !d->prt_y ||
      !d->prt_z || !d->prt_t || !d->prt_c || !d->prt_cdc || !d->prt_seq3r || !d->prt_hpp ||
      !d->prt_ele || !d->ene || !d->minus_beta || (K > 0 && (!d->kde_t || !d->kde_x)) ||
      (M > 0 && (!d->mcs_x || !d->mcs_tcc)) || (d->psp_ncol > 0 && !d->psp))
    return GD_ERR_ARG;
  const long long R64 = (long long)nl * np * nt;
  if (R64 > (1 << 26) || L > 4096 || P > (1 << 22)) return GD_ERR_SIZE;
  // the pair kernel addresses a protein conformer's points by a 32-bit byte offset
  if ((long long)np * P * 16 > 0xFFFFFFFFLL) {
    set_error("protein conformers x points over 2^28");
    return GD_ERR_SIZE;
  }
  for (int l = 0; l < L; ++l)
    if (d->lig_t[l] < 0 || d->lig_t[l] >= GD_NLT) {
      set_error("ligand atom type outside 0..23 (BADKDE)");
      return GD_ERR_TYPE;
    }
  for (int p = 0; p < P; ++p)
    if (d->prt_t[p] < 0 || d->prt_t[p] >= GD_NPT) {
      set_error("protein point type outside 0..30");
      return GD_ERR_TYPE;
    }
  const int R = (int)R64, Rp = (R + 63) / 64 * 64;
  const gd_enepara *e = d->ene;
  GdJob g;
  std::memset(&g, 0, sizeof g);
  g.L = L, g.P = P, g.K = K, g.M = M, g.n_lig = nl, g.n_prt = np, g.n_tmp = nt, g.R = R, g.Rp = Rp;

  // distinct ligand types of the job -> parameter-table rows
  std::vector<int> types;
  std::vector<int> trow(GD_NLT, -1);
  for (int l = 0; l < L; ++l)
    if (trow[d->lig_t[l]] < 0) {
      trow[d->lig_t[l]] = (int)types.size();
      types.push_back(d->lig_t[l]);
    }
  const int nty = (int)types.size();
  g.n_types = nty;
  // atom stride padded to 16 (the widest ligand chunk): the pair kernel runs whole
  // chunks; padding atoms use row block nty, all zero with a NaN psp threshold
  // (no psp, no H-bond), charge 0, no KDE bins and no MCS entries
  const int Ls = (L + 15) / 16 * 16;
  g.Ls = Ls;
  std::vector<float> ptab((size_t)(nty + 1) * GD_PTAB_PT * GD_PTAB_W, 0.0f);
  for (int pt = 0; pt < GD_PTAB_PT; ++pt) {
    ptab[((size_t)nty * GD_PTAB_PT + pt) * GD_PTAB_W + 0] = std::numeric_limits<float>::infinity();
    ptab[((size_t)nty * GD_PTAB_PT + pt) * GD_PTAB_W + 5] = std::numeric_limits<float>::quiet_NaN();
  }
  const double log2e = 1.44269504088896340736;
  const double hdb_c = std::sqrt(0.5 * log2e);
  for (int ti = 0; ti < nty; ++ti) {
    const int lt = types[ti];
    for (int pt = 0; pt < GD_NPT; ++pt) {
      float *o = &ptab[((size_t)ti * GD_PTAB_PT + pt) * GD_PTAB_W];
      const double pmf0 = e->pmf0[lt][pt];
      // LJ over p1a: w = 1 / p1a and -q = -p2a / p1a; a row without LJ
      // (p1a = p2a = 0) has w = inf, which makes the term 0
      const double p1a = e->p1a[lt][pt], p2a = e->p2a[lt][pt];
      if (p1a == 0.0 && p2a != 0.0) {
        set_error("LJ parameters with p1a = 0 and p2a != 0 are not supported");
        return GD_ERR_ARG;
      }
      o[0] = p1a != 0.0 ? (float)(1.0 / p1a) : std::numeric_limits<float>::infinity();
      o[1] = p1a != 0.0 ? (float)(-p2a / p1a) : 0.0f;
      // negated PMF exponent -(6 - r/2)(r - pmf0) log2(e) = log2(e)/2 r^2 + B r + C
      o[2] = (float)(-log2e * (6.0 + 0.5 * pmf0));
      o[3] = (float)(6.0 * log2e * pmf0);
      o[4] = e->pmf1[lt][pt];
      // psp cutoff on r^2 (compared against |o[5]|; NaN: never); the sign bit
      // marks the pairs whose H-bond term is on
      const float thr = sqrt_threshold(e->pmf0[lt][pt]);
      const bool hb = e->hdb0[lt][pt] > 0.1f;
      o[5] = std::copysign(thr >= 0.0f ? thr : std::numeric_limits<float>::quiet_NaN(), hb ? -1.0f : 1.0f);
      // H-bond: hdb1 exp(-(r - hdb0)^2 hdb1^2 / 2), the raw parameters (the
      // kernel forms the exponent in FP64); rows with hdb0 <= 0.1 are off.  The
      // near threshold covers q = ((r - hdb0) c)^2 < 160, c = hdb1 sqrt(log2(e)/2)
      const double c = (double)e->hdb1[lt][pt] * hdb_c;
      o[6] = hb ? e->hdb1[lt][pt] : 0.0f;
      o[7] = hb ? e->hdb0[lt][pt] : 0.0f;
      o[8] = near_r2_threshold(pmf0, e->el1, thr, hb, e->hdb0[lt][pt], hb ? c : 0.0);
    }
  }
  // Protein points, kept in source order: the device order depends on the tile
  // count and is laid out by layout_points (called from apply_tiling).
  // Classes: cdc flag + 2 x class-2 flag.
  {
    int nc[4] = {0, 0, 0, 0};
    auto cls = [&](int p) { return (d->prt_cdc[p] == 1 ? 1 : 0) + (d->prt_c[p] == 2 ? 2 : 0); };
    for (int p = 0; p < P; ++p) ++nc[cls(p)];
    g.pcls[0] = 0;
    for (int c = 0; c < 4; ++c) g.pcls[c + 1] = g.pcls[c] + nc[c];
    j->src_key.resize(P);
    for (int p = 0; p < P; ++p) j->src_key[p] = cls(p) * GD_PTAB_PT + d->prt_t[p];
    // within a (class, type) group: the bit-reversed rank of the distance to the
    // pocket centre (protein conformer 0), a low-discrepancy order in which any
    // run of consecutive points spans near and far points evenly
    std::vector<float> dist(P);
    for (int p = 0; p < P; ++p) {
      const float dx = d->prt_x[p] - d->pocket_center[0], dy = d->prt_y[p] - d->pocket_center[1],
                  dz = d->prt_z[p] - d->pocket_center[2];
      dist[p] = dx * dx + dy * dy + dz * dz;
    }
    std::vector<int> ord(P);
    for (int p = 0; p < P; ++p) ord[p] = p;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
      return j->src_key[x] != j->src_key[y] ? j->src_key[x] < j->src_key[y] : dist[x] < dist[y];
    });
    j->src_spread.resize(P);
    for (int i = 0; i < P;) {
      int e = i;
      while (e < P && j->src_key[ord[e]] == j->src_key[ord[i]]) ++e;
      int bits = 0;
      while ((1 << bits) < e - i) ++bits;
      for (int k = i; k < e; ++k) {
        unsigned v = (unsigned)(k - i), rv = 0;
        for (int t = 0; t < bits; ++t) rv |= ((v >> t) & 1u) << (bits - 1 - t);
        j->src_spread[ord[k]] = (int)rv;
      }
      i = e;
    }
  }
  j->src_pspv.assign((size_t)(nty + 1) * P, 0.0f);  // last row: padding atoms (zero)
  j->src_pattr.resize(P);
  for (int p = 0; p < P; ++p) {
    const int c = d->prt_c[p];
    int fl = d->prt_t[p] | (d->prt_cdc[p] == 1 ? GD_PF_CDC : 0) | (c == 2 ? GD_PF_CLS2 : 0);
    float flf;
    std::memcpy(&flf, &fl, 4);
    j->src_pattr[p] = make_float4(d->prt_ele[p], d->prt_hpp[p], flf, 0.0f);
    if (c == 2) {
      const int res = d->prt_seq3r[p];
      for (int ti = 0; ti < nty; ++ti)
        j->src_pspv[(size_t)ti * P + p] =
            (res >= 0 && res < d->psp_ncol) ? d->psp[(size_t)types[ti] * d->psp_ncol + res] : 0.0f;
    }
  }
  // KDE bins (anchors sorted by type)
  std::vector<float4> kde(std::max(1, K));
  for (int k = 0; k < K; ++k) {
    if (k > 0 && d->kde_t[k] < d->kde_t[k - 1]) {
      set_error("KDE points must be sorted by type");
      return GD_ERR_ARG;
    }
    kde[k] = make_float4(d->kde_x[k], d->kde_y[k], d->kde_z[k], 0.0f);
  }
  std::vector<GdAtom> atoms(Ls);
  for (int l = L; l < Ls; ++l) {
    GdAtom &a = atoms[l];
    std::memset(&a, 0, sizeof a);
    a.tidx = nty;
    a.hpl1 = 1.0f;
  }
  j->atom_kde.assign(L, 0);
  for (int l = 0; l < L; ++l) {
    const int lt = d->lig_t[l];
    GdAtom &a = atoms[l];
    a.tidx = trow[lt];
    a.q = d->lig_c[l];
    int b = 0;
    while (b < K && d->kde_t[b] < lt) ++b;
    int en = b;
    while (en < K && d->kde_t[en] == lt) ++en;
    a.kde_b = b;
    a.kde_e = en;
    j->atom_kde[l] = en - b;
    a.kde_inv_n = en > b ? 1.0f / (float)(en - b) : 0.0f;
    a.hpl0 = e->hpl0[lt];
    a.hpl1 = e->hpl1[lt];
    a.hpl2 = e->hpl2[lt];
  }
  std::vector<float4> mcs((size_t)std::max(1, M) * Ls, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  std::vector<float2> mcsrow(std::max(1, M));
  for (int m = 0; m < M; ++m) {
    int n = 0;
    for (int l = 0; l < L; ++l) {
      const size_t i = (size_t)m * L + l;
      const bool v = d->mcs_x[i] != GD_MCS_INVALID;
      n += v;
      if (v) mcs[(size_t)m * Ls + l] = make_float4(d->mcs_x[i], d->mcs_y[i], d->mcs_z[i], 1.0f);
    }
    mcsrow[m] = make_float2(d->mcs_tcc[m], (float)n);
  }
  std::vector<float4> lig0((size_t)nl * L);
  for (int i = 0; i < nl; ++i)
    for (int l = 0; l < L; ++l) {
      const size_t q = (size_t)i * L + l;
      lig0[q] = make_float4(d->lig_x[q], d->lig_y[q], d->lig_z[q], 0.0f);
    }
  j->src_pco.resize((size_t)np * P);
  for (size_t q = 0; q < (size_t)np * P; ++q)
    j->src_pco[q] = make_float4(d->prt_x[q], d->prt_y[q], d->prt_z[q], 0.0f);
  // pmf1 summed over every (atom, point) pair: the PMF term of a pose is this
  // constant minus the pair kernel's corrections
  {
    double sp = 0.0;
    for (int l = 0; l < L; ++l)
      for (int p = 0; p < P; ++p) sp += e->pmf1[d->lig_t[l]][d->prt_t[p]];
    g.pmf_all = (float)sp;
  }
  // scalars
  g.lj0 = e->lj0, g.lj1 = e->lj1, g.el0 = e->el0, g.el1 = e->el1, g.a1 = e->a1, g.b1 = e->b1;
  g.inv_el1 = 1.0f / e->el1;
  g.kde2l = e->kde2 * 1.44269504088896340736f;
  g.kde3_inv = 1.0f / e->kde3;
  g.inv_L = 1.0f / (float)L;
  for (int i = 0; i < 3; ++i) {
    g.center[i] = d->lig_center[i];
    g.pocket[i] = d->pocket_center[i];
    g.cp[i] = (float)((double)d->lig_center[i] - (double)d->pocket_center[i]);
  }
  for (int i = 0; i < 10; ++i) g.a_para[i] = e->a_para[i], g.b_para[i] = e->b_para[i], g.w[i] = e->w[i];

  // device memory: constant tables + replica state
  Arena a;
  float4 *d_lig0, *d_pco, *d_pattr, *d_kde, *d_mcs;
  GdAtom *d_atoms;
  float *d_ptab, *d_pspv, *d_mb;
  float2 *d_mcsrow;
  a.add(&d_lig0, lig0.size());
  a.add(&d_atoms, atoms.size());
  a.add(&d_pco, j->src_pco.size());
  a.add(&d_pattr, j->src_pattr.size());
  a.add(&d_ptab, ptab.size());
  a.add(&d_pspv, j->src_pspv.size());
  a.add(&d_kde, kde.size());
  a.add(&d_mcs, mcs.size());
  a.add(&d_mcsrow, mcsrow.size());
  a.add(&d_mb, (size_t)nt);
  a.add(&g.mm, (size_t)6 * Rp);
  a.add(&g.prop, (size_t)6 * Rp);
  a.add(&g.en, (size_t)10 * Rp);
  a.add(&g.ligi, (size_t)Rp);
  a.add(&g.prti, (size_t)Rp);
  a.add(&g.tmpi, (size_t)Rp);
  a.add(&g.acc, (size_t)Rp);
  a.add(&g.nacc, (size_t)Rp);
  a.add(&g.nexc, (size_t)Rp);
  a.add(&g.nswp, (size_t)2 * Rp);
  a.add(&g.swp_e, (size_t)Rp);
  a.add(&g.best, (size_t)Rp);
  a.add(&g.own, (size_t)nl * np * nt);
  a.add(&g.coords, (size_t)3 * L * Rp);
  a.add(&g.rec_n, (size_t)2 * Rp);
  a.add(&g.ev_terms, (size_t)10 * Rp);
  a.add(&g.ev_raw, (size_t)9 * Rp);
  a.add(&g.wq, 64);
  gd_status st = a.commit(&j->d_base);
  if (st != GD_OK) return st;
  j->bytes = a.off;
  g.lig0 = d_lig0, g.atoms = d_atoms, g.pcoord = d_pco, g.pattr = d_pattr, g.ptab = d_ptab;
  g.pspv = d_pspv, g.kde = d_kde, g.mcs = d_mcs, g.mcsrow = d_mcsrow, g.minus_beta = d_mb;
  auto up = [&](void *dst, const void *src, size_t n) -> gd_status {
    hipError_t e2 = hipMemcpy(dst, src, n, hipMemcpyHostToDevice);
    return e2 == hipSuccess ? GD_OK : hip_fail(e2, "hipMemcpy H2D");
  };
  if ((st = up(d_lig0, lig0.data(), lig0.size() * sizeof(float4))) != GD_OK) return st;
  if ((st = up(d_atoms, atoms.data(), atoms.size() * sizeof(GdAtom))) != GD_OK) return st;
  if ((st = up(d_ptab, ptab.data(), ptab.size() * sizeof(float))) != GD_OK) return st;
  if ((st = up(d_kde, kde.data(), kde.size() * sizeof(float4))) != GD_OK) return st;
  if ((st = up(d_mcs, mcs.data(), mcs.size() * sizeof(float4))) != GD_OK) return st;
  if ((st = up(d_mcsrow, mcsrow.data(), mcsrow.size() * sizeof(float2))) != GD_OK) return st;
  if ((st = up(d_mb, d->minus_beta, (size_t)nt * sizeof(float))) != GD_OK) return st;
  j->h = g;
  return GD_OK;
}

// Replica table of SetReplica (util.C:370-406): idx = n_tmp*n_lig*p + n_lig*t + l,
// zero pose and energies; padding lanes get conformer 0 and never write back.
// Device bytes gd_job_create allocates for a descriptor (tables, replica
// state, partial sums under the default tiling, work lists).
// Device bytes of a complex under the plain tiling, or (shared) under the
// tiling of shared launches.
size_t footprint(const gd_job_desc *d, bool shared) {
  if (!d || d->lig_natom < 1 || d->prt_npoint < 1 || d->n_lig < 1 || d->n_prt < 1 || d->n_tmp < 1)
    return 0;
  const size_t L = d->lig_natom, P = d->prt_npoint, K = d->kde_npoint, M = d->mcs_nrow;
  const size_t nl = d->n_lig, np = d->n_prt, nt = d->n_tmp;
  const size_t R = nl * np * nt, Rp = (R + 63) / 64 * 64;
  auto a = [](size_t bytes) { return (bytes + 255) / 256 * 256; };
  std::vector<int> seen(GD_NLT, 0);
  size_t nty = 0;
  for (size_t l = 0; l < L; ++l)
    if (d->lig_t[l] >= 0 && d->lig_t[l] < GD_NLT && !seen[d->lig_t[l]]++) ++nty;
  const size_t Ls = (L + 15) / 16 * 16;
  size_t b = a(nl * L * 16) + a(Ls * sizeof(GdAtom)) + a(np * P * 16) + a(P * 16) +
             a((nty + 1) * GD_PTAB_PT * GD_PTAB_W * 4) + a((nty + 1) * P * 4) +
             a(std::max<size_t>(1, K) * 16) + a(std::max<size_t>(1, M) * Ls * 16) +
             a(std::max<size_t>(1, M) * 8) + a(nt * 4);
  b += a(6 * Rp * 4) * 2 + a(10 * Rp * 4) + a(Rp * 4) * 6 + a(2 * Rp * 4) + a(Rp * 4) +
       a(Rp * 4) + a(nl * np * nt * 4) +
       a(3 * L * Rp * 4) + a(2 * Rp * 4) + a(10 * Rp * 4) + a(9 * Rp * 4) + a(64 * 4);
  int lc, ptile;
  if (shared) {
    lc = 4;
    ptile = shared_ptile((int)L, (int)P);
  } else {
    default_tiling((int)L, (int)P, (int)R, lc, ptile);
  }
  const size_t nL = (L + lc - 1) / lc, nP0 = (P + ptile - 1) / ptile;
  const size_t nG = (nP0 + GD_GROUP - 1) / GD_GROUP;
  b += a(GD_NLIN * nL * nG * Rp * 4) + a(L * nG * Rp * 4) + a(std::max<size_t>(1, M) * nL * Rp * 4) +
       a((Rp / 64) * GD_RK * 8 * 64 * 4);
  b += (Rp / 64) * nL * nG * sizeof(GdItem) + (Rp / 64) * sizeof(GdBlock) + 512;
  return b;
}

gd_status reset_replicas(gd_job *j) {
  const GdJob &g = j->h;
  const int Rp = g.Rp, nl = g.n_lig, np = g.n_prt, nt = g.n_tmp;
  std::vector<int> ligi(Rp, 0), prti(Rp, 0), tmpi(Rp, 0), zero(Rp, 0), own((size_t)nl * np * nt);
  for (int p = 0; p < np; ++p)
    for (int t = 0; t < nt; ++t)
      for (int l = 0; l < nl; ++l) {
        const int r = nt * nl * p + nl * t + l;
        ligi[r] = l, prti[r] = p, tmpi[r] = t;
        own[((size_t)p * nl + l) * nt + t] = r;
      }
  std::vector<float> zf((size_t)10 * Rp, 0.0f), best(Rp, 1e30f);
  hipStream_t s = j->stream;
  hipError_t e;
#define GD_UP(dst, src, n)                                                          \
  if ((e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s)) != hipSuccess) \
    return hip_fail(e, "hipMemcpyAsync");
  GD_UP(g.ligi, ligi.data(), Rp * sizeof(int));
  GD_UP(g.prti, prti.data(), Rp * sizeof(int));
  GD_UP(g.tmpi, tmpi.data(), Rp * sizeof(int));
  GD_UP(g.acc, zero.data(), Rp * sizeof(int));
  GD_UP(g.nacc, zero.data(), Rp * sizeof(int));
  GD_UP(g.nexc, zero.data(), Rp * sizeof(int));
  GD_UP(g.nswp, zero.data(), Rp * sizeof(int));
  GD_UP(g.nswp + Rp, zero.data(), Rp * sizeof(int));
  GD_UP(g.rec_n, zero.data(), Rp * sizeof(int));
  GD_UP(g.rec_n + Rp, zero.data(), Rp * sizeof(int));
  GD_UP(g.mm, zf.data(), (size_t)6 * Rp * sizeof(float));
  GD_UP(g.prop, zf.data(), (size_t)6 * Rp * sizeof(float));
  GD_UP(g.en, zf.data(), (size_t)10 * Rp * sizeof(float));
  GD_UP(g.best, best.data(), Rp * sizeof(float));
  GD_UP(g.own, own.data(), own.size() * sizeof(int));
#undef GD_UP
  e = hipStreamSynchronize(s);
  return e == hipSuccess ? GD_OK : hip_fail(e, "hipStreamSynchronize");
}

gd_status sync_job_struct(gd_job *j) {
  hipError_t e = hipMemcpyAsync(j->d_job, &j->h, sizeof(GdJob), hipMemcpyHostToDevice, j->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync job");
  std::vector<GdBlock> blocks;
  for (int rb = 0; rb < j->h.Rp; rb += 64) blocks.push_back({0, rb});
  j->n_blocks = (int)blocks.size();
  j->n_items = (int)j->items.size();
  // the plain path's stepping parts: n_streams contiguous ranges of the
  // blocks, as even as the block count allows, each with the job's items of
  // its blocks (in the job's order)
  const int np = std::max(1, std::min(j->n_streams, j->n_blocks));
  j->n_parts = np;
  std::vector<GdItem> part_items;
  for (int k = 0; k < np; ++k) {
    j->b0_part[k] = (int)((long)j->n_blocks * k / np);
    j->nb_part[k] = (int)((long)j->n_blocks * (k + 1) / np) - j->b0_part[k];
    const int lo = j->b0_part[k] * 64, hi = (j->b0_part[k] + j->nb_part[k]) * 64;
    j->n_items_part[k] = 0;
    for (const GdItem &it : j->items)
      if (it.rb >= lo && it.rb < hi) {
        part_items.push_back(it);
        ++j->n_items_part[k];
      }
  }
  if (j->d_lists) hipFree(j->d_lists);
  j->d_lists = nullptr;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t bi = (size_t)j->n_items * sizeof(GdItem), bb = (size_t)j->n_blocks * sizeof(GdBlock),
               bh = part_items.size() * sizeof(GdItem);
  j->lists_bytes = 0;
  e = hipMalloc(&j->d_lists, up(bi) + up(bb) + up(bh) + 256);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc lists");
  j->lists_bytes = up(bi) + up(bb) + up(bh) + 256;
  j->d_items = (GdItem *)j->d_lists;
  j->d_blocks = (GdBlock *)((char *)j->d_lists + up(bi));
  j->d_items_part[0] = (GdItem *)((char *)j->d_lists + up(bi) + up(bb));
  for (int k = 1; k < np; ++k) j->d_items_part[k] = j->d_items_part[k - 1] + j->n_items_part[k - 1];
  if ((e = hipMemcpyAsync(j->d_items, j->items.data(), bi, hipMemcpyHostToDevice, j->stream)) != hipSuccess)
    return hip_fail(e, "hipMemcpyAsync items");
  if ((e = hipMemcpyAsync(j->d_blocks, blocks.data(), bb, hipMemcpyHostToDevice, j->stream)) != hipSuccess)
    return hip_fail(e, "hipMemcpyAsync blocks");
  if (bh && (e = hipMemcpyAsync(j->d_items_part[0], part_items.data(), bh, hipMemcpyHostToDevice,
                                j->stream)) != hipSuccess)
    return hip_fail(e, "hipMemcpyAsync part items");
  e = hipStreamSynchronize(j->stream);
  return e == hipSuccess ? GD_OK : hip_fail(e, "hipStreamSynchronize");
}

// ---- launches ---------------------------------------------------------------
gd_status launch_pair(int lc, const GdJob *d_jobs, const GdItem *d_items, int n_items,
                      hipStream_t s) {
  if (n_items <= 0) return GD_OK;
  const dim3 grid(n_items), block(64 * GD_GROUP);  // one group per item
  switch (lc) {
    case 4:
      hipLaunchKernelGGL(gd_pair_kernel<4>, grid, block, 0, s, d_jobs, d_items, n_items);
      break;
    case 8:
      hipLaunchKernelGGL(gd_pair_kernel<8>, grid, block, 0, s, d_jobs, d_items, n_items);
      break;
    case 16:
      hipLaunchKernelGGL(gd_pair_kernel<16>, grid, block, 0, s, d_jobs, d_items, n_items);
      break;
    default:
      return GD_ERR_ARG;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GD_OK : hip_fail(e, "pair kernel launch");
}

gd_status launch_finish(const GdJob *d_jobs, const GdBlock *d_blocks, int n_blocks, int mode,
                        int step, hipStream_t s) {
  if (n_blocks <= 0) return GD_OK;
  hipLaunchKernelGGL(gd_reduce_kernel, dim3(n_blocks * GD_RK), dim3(256), 0, s, d_jobs, d_blocks);
  hipLaunchKernelGGL(gd_finish_kernel, dim3(n_blocks), dim3(1024), 0, s, d_jobs, d_blocks, mode, step,
                     (const int *)nullptr, (int *)nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GD_OK : hip_fail(e, "finish kernel launch");
}

gd_status launch_transform(const GdJob *d_jobs, const GdBlock *d_blocks, int n_blocks,
                           hipStream_t s) {
  if (n_blocks <= 0) return GD_OK;
  hipLaunchKernelGGL(gd_transform_kernel, dim3(n_blocks), dim3(256), 0, s, d_jobs, d_blocks,
                     (int *)nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GD_OK : hip_fail(e, "transform kernel launch");
}

gd_status launch_exchange(const GdJob *d_jobs, const GdExch *d_ex, int n_ex, hipStream_t s) {
  if (n_ex <= 0) return GD_OK;
  hipLaunchKernelGGL(gd_exchange_kernel, dim3(n_ex), dim3(256), 0, s, d_jobs, d_ex);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GD_OK : hip_fail(e, "exchange kernel launch");
}

gd_status launch_swap(int axis, const GdJob *d_jobs, const GdBlock *d_blocks, int n_blocks,
                      const int *d_swr, int launch_step, const int *step_base, int *wq,
                      const std::function<gd_status()> &pair, hipStream_t s) {
  if (n_blocks <= 0) return GD_OK;
  hipLaunchKernelGGL(gd_swap_propose_kernel, dim3(n_blocks), dim3(256), 0, s, d_jobs, d_blocks, axis,
                     d_swr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "swap propose kernel launch");
  gd_status st = pair();
  if (st != GD_OK) return st;
  hipLaunchKernelGGL(gd_reduce_kernel, dim3(n_blocks * GD_RK), dim3(256), 0, s, d_jobs, d_blocks);
  hipLaunchKernelGGL(gd_finish_kernel, dim3(n_blocks), dim3(1024), 0, s, d_jobs, d_blocks, GD_MODE_EVAL,
                     0, (const int *)nullptr, wq);
  hipLaunchKernelGGL(gd_swap_accept_kernel, dim3(n_blocks), dim3(256), 0, s, d_jobs, d_blocks, axis,
                     d_swr, launch_step, step_base);
  e = hipGetLastError();
  return e == hipSuccess ? GD_OK : hip_fail(e, "swap kernel launch");
}

void set_mc(GdJob &g, const gd_mc_param *mc) {
  for (int i = 0; i < 6; ++i) g.move_scale[i] = mc->move_scale[i];
  g.move_mode = mc->move_mode;
  g.seed_lo = (uint32_t)mc->seed;
  g.seed_hi = (uint32_t)(mc->seed >> 32);
  g.job_id = mc->job_id;
  g.record = mc->record;
  g.xaxes = mc->exchange_axes ? mc->exchange_axes : GD_EXCH_TEMP;
  g.steps_total = mc->steps_total;
  g.step0 = 0;
}

gd_status read_out(gd_job *j, gd_replica_out *out) {
  if (!out) return GD_OK;
  const GdJob &g = j->h;
  const int R = g.R, Rp = g.Rp;
  hipError_t e;
  std::vector<float> f((size_t)10 * Rp);
  std::vector<int> iv(Rp);
  // copies on the job's stream: ordered after every kernel of the run
  auto down = [&](void *dst, const void *src, size_t n) -> gd_status {
    e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, j->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    return e == hipSuccess ? GD_OK : hip_fail(e, "hipMemcpy D2H");
  };
  gd_status st;
  if (out->movematrix) {
    if ((st = down(f.data(), g.mm, (size_t)6 * Rp * sizeof(float))) != GD_OK) return st;
    for (int r = 0; r < R; ++r)
      for (int i = 0; i < 6; ++i) out->movematrix[(size_t)r * 6 + i] = f[(size_t)i * Rp + r];
  }
  if (out->energy) {
    if ((st = down(f.data(), g.en, (size_t)10 * Rp * sizeof(float))) != GD_OK) return st;
    for (int r = 0; r < R; ++r)
      for (int i = 0; i < 10; ++i) out->energy[(size_t)r * 10 + i] = f[(size_t)i * Rp + r];
  }
  if (out->best_energy) {
    if ((st = down(f.data(), g.best, (size_t)Rp * sizeof(float))) != GD_OK) return st;
    std::memcpy(out->best_energy, f.data(), (size_t)R * sizeof(float));
  }
  struct {
    int *dst;
    int *src;
  } ints[] = {{out->idx_tmp, g.tmpi}, {out->n_accept, g.nacc}, {out->n_exchange, g.nexc},
              {out->is_accept, g.acc}, {out->n_swap_lig, g.nswp}, {out->n_swap_prt, g.nswp + Rp}};
  for (auto &q : ints)
    if (q.dst) {
      if ((st = down(iv.data(), q.src, (size_t)Rp * sizeof(int))) != GD_OK) return st;
      std::memcpy(q.dst, iv.data(), (size_t)R * sizeof(int));
    }
  return GD_OK;
}

}  // namespace gdr

using namespace gdr;

// ---------------------------------------------------------------------------
extern "C" {

gd_status gd_init(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    set_error("no HIP device visible");
    return GD_ERR_NODEV;
  }
  if (dev < 0 || dev >= n) return GD_ERR_ARG;
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
    return GD_ERR_NODEV;
  }
  if ((e = hipSetDevice(dev)) != hipSuccess) return hip_fail(e, "hipSetDevice");
  g_device = dev;
  g_cu = prop.multiProcessorCount;
  return GD_OK;
}

gd_status gd_device_info(char *name, int name_len, int *n_cu_out, int *clock_khz, size_t *hbm) {
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, g_device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (name && name_len > 0) {
    if (prop.name[0])
      std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    else  // the runtime reports no marketing name: the architecture alone
      std::snprintf(name, name_len, "%s, %d CUs", prop.gcnArchName, prop.multiProcessorCount);
  }
  if (n_cu_out) *n_cu_out = prop.multiProcessorCount;
  if (clock_khz) *clock_khz = prop.clockRate;
  if (hbm) *hbm = prop.totalGlobalMem;
  return GD_OK;
}

const char *gd_last_error(void) { return last_error(); }
const char *gd_version(void) { return "geauxdock_amd 0.1 (gfx950)"; }

// The pair kernel's far-field r^2 threshold for (pmf0, el1) (diagnostics, tests).
float gd_far_r2(float pmf0, float el1) { return gdr::far_r2_threshold(pmf0, el1); }

// The row threshold of the pair loop's near branch for a (ligand type, point
// type) pair of the force field (diagnostics, tests): as the device row holds it.
float gd_near_r2(float pmf0, float el1, float hdb0, float hdb1) {
  const bool hb = hdb0 > 0.1f;
  const double c = (double)hdb1 * std::sqrt(0.5 * 1.44269504088896340736);
  return gdr::near_r2_threshold(pmf0, el1, gdr::sqrt_threshold(pmf0), hb, hdb0, hb ? c : 0.0);
}

void gd_philox(const uint32_t ctr[4], const uint32_t key[2], int rounds, uint32_t out[4]) {
  const gd_u32x4 r = gd_philox4x32(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], rounds);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

gd_status gd_job_create(const gd_job_desc *desc, gd_job **out) {
  if (!desc || !out) return GD_ERR_ARG;
  *out = nullptr;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  gd_job *j = new gd_job();
  hipError_t e = hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete j;
    return hip_fail(e, "hipStreamCreate");
  }
  st = pack_job(desc, j);
  if (st == GD_OK) {
    e = hipMalloc(&j->d_job, sizeof(GdJob));
    if (e != hipSuccess) st = hip_fail(e, "hipMalloc job");
  }
  if (st == GD_OK) {
    int lc, ptile;
    default_tiling(j->h.L, j->h.P, j->h.R, lc, ptile);
    st = apply_tiling(j, lc, ptile);
  }
  if (st == GD_OK) st = sync_job_struct(j);
  if (st != GD_OK) {
    gd_job_destroy(j);
    return st;
  }
  *out = j;
  return GD_OK;
}

gd_status gd_job_destroy(gd_job *j) {
  if (!j) return GD_OK;
  if (j->d_base) hipFree(j->d_base);
  if (j->d_part) hipFree(j->d_part);
  if (j->d_job) hipFree(j->d_job);
  if (j->d_lists) hipFree(j->d_lists);
  if (j->d_rec) hipFree(j->d_rec);
  if (j->h_rec) hipHostFree(j->h_rec);
  if (j->copy_stream) hipStreamDestroy(j->copy_stream);
  for (hipStream_t ps : j->part_stream)
    if (ps) hipStreamDestroy(ps);
  if (j->stream) hipStreamDestroy(j->stream);
  delete j;
  return GD_OK;
}

gd_status gd_job_sizes(const gd_job *j, int *n_rep, int *L, int *P, int *K, int *M) {
  if (!j) return GD_ERR_ARG;
  if (n_rep) *n_rep = j->h.R;
  if (L) *L = j->h.L;
  if (P) *P = j->h.P;
  if (K) *K = j->h.K;
  if (M) *M = j->h.M;
  return GD_OK;
}

size_t gd_job_footprint(const gd_job_desc *d) { return footprint(d, false); }

// Device memory the job holds: tables and replica state, partial sums, work
// lists, the record ring (when a run recorded) and the device job struct.
size_t gd_job_device_bytes(const gd_job *j) {
  if (!j) return 0;
  return j->bytes + j->part_bytes + j->lists_bytes + j->rec_bytes + sizeof(GdJob);
}

gd_status gd_tiling_get(const gd_job *j, gd_tiling *t) {
  if (!j || !t) return GD_ERR_ARG;
  t->atoms_per_lane = j->h.lc;
  t->tile_points = j->h.ptile;
  t->waves_per_cu = 0;
  return GD_OK;
}

gd_status gd_job_set_streams(gd_job *j, int n_streams) {
  if (!j || n_streams < 1 || n_streams > gd_job::kMaxParts) return GD_ERR_ARG;
  if (n_streams == j->n_streams) return GD_OK;
  const gd_status st = ensure_device();
  if (st != GD_OK) return st;
  j->n_streams = n_streams;
  return sync_job_struct(j);  // the parts' item lists
}

gd_status gd_tiling_set(gd_job *j, const gd_tiling *t) {
  if (!j || !t) return GD_ERR_ARG;
  gd_status st = apply_tiling(j, t->atoms_per_lane, t->tile_points);
  if (st != GD_OK) return st;
  return sync_job_struct(j);
}

// Energies of arbitrary poses through the same pair/finish kernels: the poses
// are loaded as a temporary replica set (chunks of at most the job's padded
// replica count), transformed, evaluated in EVAL mode, and the job's own
// replica state is restored afterwards.
gd_status gd_energy_eval(gd_job *j, int n, const int *conf, const float *pose, float *terms,
                         float *raw, float *coords) {
  if (!j || n < 0 || (n > 0 && (!conf || !pose || !terms))) return GD_ERR_ARG;
  if (n == 0) return GD_OK;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  GdJob &g = j->h;
  for (int i = 0; i < n; ++i)
    if (conf[2 * i] < 0 || conf[2 * i] >= g.n_lig || conf[2 * i + 1] < 0 || conf[2 * i + 1] >= g.n_prt)
      return GD_ERR_ARG;
  const int Rp = g.Rp, L = g.L;
  hipStream_t s = j->stream;
  // save the replica conformer maps, restore at the end
  std::vector<int> ligi_save(Rp), prti_save(Rp);
  std::vector<float> prop_save((size_t)6 * Rp);
  hipError_t e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess ||
      (e = hipMemcpy(ligi_save.data(), g.ligi, Rp * sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess ||
      (e = hipMemcpy(prti_save.data(), g.prti, Rp * sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess ||
      (e = hipMemcpy(prop_save.data(), g.prop, (size_t)6 * Rp * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "hipMemcpy save");
  std::vector<int> li(Rp), pi(Rp);
  std::vector<float> pp((size_t)6 * Rp), ot((size_t)10 * Rp), orw((size_t)9 * Rp),
      oc((size_t)3 * L * Rp);
  const int saveR = g.R;
  for (int base = 0; base < n && st == GD_OK; base += Rp) {
    const int cnt = std::min(Rp, n - base);
    std::fill(li.begin(), li.end(), 0);
    std::fill(pi.begin(), pi.end(), 0);
    std::fill(pp.begin(), pp.end(), 0.0f);
    for (int i = 0; i < cnt; ++i) {
      li[i] = conf[2 * (base + i)];
      pi[i] = conf[2 * (base + i) + 1];
      for (int k = 0; k < 6; ++k) pp[(size_t)k * Rp + i] = pose[6 * (size_t)(base + i) + k];
    }
    g.R = cnt;
    if ((e = hipMemcpyAsync(j->d_job, &g, sizeof(GdJob), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(g.ligi, li.data(), Rp * sizeof(int), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(g.prti, pi.data(), Rp * sizeof(int), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(g.prop, pp.data(), (size_t)6 * Rp * sizeof(float), hipMemcpyHostToDevice, s)) != hipSuccess) {
      st = hip_fail(e, "hipMemcpyAsync eval");
      break;
    }
    if ((st = launch_transform(j->d_job, j->d_blocks, j->n_blocks, s)) != GD_OK) break;
    if ((st = launch_pair(g.lc, j->d_job, j->d_items, j->n_items, s)) != GD_OK) break;
    if ((st = launch_finish(j->d_job, j->d_blocks, j->n_blocks, GD_MODE_EVAL, 0, s)) != GD_OK) break;
    if ((e = hipMemcpyAsync(ot.data(), g.ev_terms, (size_t)10 * Rp * sizeof(float), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(orw.data(), g.ev_raw, (size_t)9 * Rp * sizeof(float), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (coords && (e = hipMemcpyAsync(oc.data(), g.coords, (size_t)3 * L * Rp * sizeof(float), hipMemcpyDeviceToHost, s)) != hipSuccess) ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
      st = hip_fail(e, "eval readback");
      break;
    }
    for (int i = 0; i < cnt; ++i) {
      for (int k = 0; k < 10; ++k) terms[(size_t)(base + i) * 10 + k] = ot[(size_t)i * 10 + k];
      if (raw)
        for (int k = 0; k < 9; ++k) raw[(size_t)(base + i) * 9 + k] = orw[(size_t)i * 9 + k];
      if (coords)
        for (int l = 0; l < L; ++l)
          for (int d = 0; d < 3; ++d)
            coords[((size_t)(base + i) * L + l) * 3 + d] = oc[(size_t)(3 * l + d) * Rp + i];
    }
  }
  g.R = saveR;
  hipMemcpyAsync(g.ligi, ligi_save.data(), Rp * sizeof(int), hipMemcpyHostToDevice, s);
  hipMemcpyAsync(g.prti, prti_save.data(), Rp * sizeof(int), hipMemcpyHostToDevice, s);
  hipMemcpyAsync(g.prop, prop_save.data(), (size_t)6 * Rp * sizeof(float), hipMemcpyHostToDevice, s);
  hipMemcpyAsync(j->d_job, &g, sizeof(GdJob), hipMemcpyHostToDevice, s);
  e = hipStreamSynchronize(s);
  if (st == GD_OK && e != hipSuccess) st = hip_fail(e, "eval restore");
  return st;
}

// Plain one-job docking (the reference Dock shape, backend_gpu.cu:64-133):
// initial-energy pass, then steps_total steps in dump segments; temperature
// exchange every steps_per_exchange steps; with mc->record the accept records
// of the last segment stay on the device for gd_dock_records.
// Receiver of the per-dump accept records (gd_dock_stream); null: the records
// stay on the device (gd_dock_records reads the last dump).
struct RecSink {
  gd_record_fn fn;
  void *user;
  bool blocking;
};

// The plain driver: initial-energy pass, then steps_total MC steps in dumps of
// steps_per_dump (Dock, backend_gpu.cu:64-133), exchange points every
// steps_per_exchange steps.  With a sink, dump k's accept records (ring half
// k mod 2) go to the host after the dump's kernels: blocking, as the
// reference's cudaMemcpy after each dump (backend_gpu.cu:110-116), or on the
// copy stream while the device runs dump k + 1, handed to the sink once
// copied; dump k + 2 reuses the half only after the host has taken it.
static gd_status dock_impl(gd_job *j, const gd_mc_param *mc, const RecSink *sink,
                           gd_replica_out *out, gd_stats *stats) {
  if (!j || !mc || mc->steps_total < 0 || mc->steps_per_exchange < 0) return GD_ERR_ARG;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  GdJob &g = j->h;
  hipStream_t s = j->stream;
  set_mc(g, mc);
  const bool record = mc->record || sink;
  g.record = record ? 1 : 0;
  const int spd = mc->steps_per_dump > 0 ? mc->steps_per_dump : std::max(1, mc->steps_total);
  const size_t half_floats = (size_t)g.Rp * spd * GD_REC_W;
  j->last_half = -1;
  if (record) {
    const size_t need = 2 * half_floats * sizeof(float);
    if (need > j->rec_bytes) {
      if (j->d_rec) hipFree(j->d_rec);
      j->d_rec = nullptr;
      j->rec_bytes = 0;
      hipError_t e = hipMalloc(&j->d_rec, need);
      if (e != hipSuccess) return hip_fail(e, "hipMalloc records");
      j->rec_bytes = need;
    }
    g.rec = (float *)j->d_rec;
    g.rec_cap = spd;
  } else {
    g.rec = nullptr;
    g.rec_cap = 0;
  }
  // host side of a streamed run: pinned halves (records, then counts) and the copy stream
  float *h_half[2] = {nullptr, nullptr};
  int *h_cnt[2] = {nullptr, nullptr};
  hipEvent_t ev_done[2] = {nullptr, nullptr}, ev_copied[2] = {nullptr, nullptr};
  if (sink) {
    const size_t hb = 2 * (half_floats * sizeof(float) + (size_t)g.Rp * sizeof(int));
    if (hb > j->h_rec_bytes) {
      if (j->h_rec) hipHostFree(j->h_rec);
      j->h_rec = nullptr;
      j->h_rec_bytes = 0;
      hipError_t e = hipHostMalloc(&j->h_rec, hb, hipHostMallocDefault);
      if (e != hipSuccess) return hip_fail(e, "hipHostMalloc records");
      j->h_rec_bytes = hb;
    }
    for (int b = 0; b < 2; ++b) {
      h_half[b] = (float *)j->h_rec + b * half_floats;
      h_cnt[b] = (int *)((float *)j->h_rec + 2 * half_floats) + (size_t)b * g.Rp;
    }
    if (!sink->blocking) {
      hipError_t e = hipSuccess;
      if (!j->copy_stream) e = hipStreamCreateWithFlags(&j->copy_stream, hipStreamNonBlocking);
      for (int b = 0; b < 2 && e == hipSuccess; ++b) {
        e = hipEventCreateWithFlags(&ev_done[b], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_copied[b], hipEventDisableTiming);
      }
      if (e != hipSuccess) {
        for (int b = 0; b < 2; ++b) {
          if (ev_done[b]) hipEventDestroy(ev_done[b]);
          if (ev_copied[b]) hipEventDestroy(ev_copied[b]);
        }
        return hip_fail(e, "record stream set-up");
      }
    }
  }
  struct EvGuard {
    hipEvent_t *a, *b;
    ~EvGuard() {
      for (int k = 0; k < 2; ++k) {
        if (a[k]) hipEventDestroy(a[k]);
        if (b[k]) hipEventDestroy(b[k]);
      }
    }
  } ev_guard{ev_done, ev_copied};
  if ((st = reset_replicas(j)) != GD_OK) return st;
  hipError_t e = hipMemcpyAsync(j->d_job, &g, sizeof(GdJob), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync job");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  // exchange points: after MC steps spe-1, 2spe-1, ...; round k uses table entry k
  const int spe = mc->steps_per_exchange;
  const bool tx = spe > 0 && (g.xaxes & GD_EXCH_TEMP) && g.n_tmp > 1;
  const bool lx = spe > 0 && (g.xaxes & GD_EXCH_LIG) && g.n_lig > 1;
  const bool px = spe > 0 && (g.xaxes & GD_EXCH_PRT) && g.n_prt > 1;
  const int n_rounds = (tx || lx || px) ? mc->steps_total / spe : 0;
  void *d_xt = nullptr;  // GdExch [n_rounds], then the swap rounds int [n_rounds]
  GdExch *d_ex = nullptr;
  int *d_swr = nullptr;
  if (n_rounds > 0) {
    const size_t be = ((size_t)n_rounds * sizeof(GdExch) + 255) / 256 * 256;
    std::vector<char> xt(be + (size_t)n_rounds * sizeof(int));
    for (int k = 0; k < n_rounds; ++k) {
      const GdExch x = {0, k};
      std::memcpy(xt.data() + k * sizeof(GdExch), &x, sizeof x);
      std::memcpy(xt.data() + be + k * sizeof(int), &k, sizeof k);
    }
    e = hipMalloc(&d_xt, xt.size());
    if (e != hipSuccess) return hip_fail(e, "hipMalloc exch");
    e = hipMemcpy(d_xt, xt.data(), xt.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipFree(d_xt);
      return hip_fail(e, "hipMemcpy exch");
    }
    d_ex = (GdExch *)d_xt;
    d_swr = (int *)((char *)d_xt + be);
  }
  const std::function<gd_status()> pair = [&]() {
    return launch_pair(g.lc, j->d_job, j->d_items, j->n_items, s);
  };
  // hand dump k (ring half b, copied to the host) to the sink
  const auto deliver = [&](int k, int b) {
    sink->fn(sink->user, k, k * spd, g.R, spd, h_half[b], h_cnt[b]);
  };
  KTimer tp;
  SpanTimer tspan;
  const bool timing = stats != nullptr;
  // Several streams, each stepping one contiguous part of the replica blocks:
  // one part's reduce and finish kernels and its launch tail overlap the other
  // parts' pair work.  Between exchange points the parts share nothing; they
  // join for every exchange point (which spans the whole job) and at dump ends.
  const int np = j->n_parts;
  bool multi = np >= 2;
  for (int k = 0; k < np; ++k) multi = multi && j->nb_part[k] > 0;
  hipStream_t ps[gd_job::kMaxParts] = {s};
  hipEvent_t ev_fork = nullptr, ev_join[gd_job::kMaxParts] = {};
  struct ForkGuard {
    hipEvent_t *f, *jn;
    ~ForkGuard() {
      if (*f) hipEventDestroy(*f);
      for (int k = 0; k < gd_job::kMaxParts; ++k)
        if (jn[k]) hipEventDestroy(jn[k]);
    }
  } fork_guard{&ev_fork, ev_join};
  if (multi) {
    e = hipSuccess;
    for (int k = 1; k < np && e == hipSuccess; ++k) {
      if (!j->part_stream[k]) e = hipStreamCreateWithFlags(&j->part_stream[k], hipStreamNonBlocking);
      ps[k] = j->part_stream[k];
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming);
    for (int k = 1; k < np && e == hipSuccess; ++k)
      e = hipEventCreateWithFlags(&ev_join[k], hipEventDisableTiming);
    if (e != hipSuccess) {
      if (d_xt) hipFree(d_xt);
      return hip_fail(e, "part streams");
    }
  }
  // the other parts' streams continue after everything queued on s so far /
  // s continues after everything queued on the other parts' streams
  const auto fork = [&]() {
    hipError_t r = hipEventRecord(ev_fork, s);
    for (int k = 1; k < np && r == hipSuccess; ++k) r = hipStreamWaitEvent(ps[k], ev_fork, 0);
    return r == hipSuccess ? GD_OK : hip_fail(r, "stream fork");
  };
  const auto join = [&]() {
    hipError_t r = hipSuccess;
    for (int k = 1; k < np && r == hipSuccess; ++k) {
      r = hipEventRecord(ev_join[k], ps[k]);
      if (r == hipSuccess) r = hipStreamWaitEvent(s, ev_join[k], 0);
    }
    return r == hipSuccess ? GD_OK : hip_fail(r, "stream join");
  };
  const auto t0 = std::chrono::steady_clock::now();
  // initial energies (MonteCarlo_d with s2max = 1)
  if (st == GD_OK) st = launch_transform(j->d_job, j->d_blocks, j->n_blocks, s);
  if (st == GD_OK) st = launch_pair(g.lc, j->d_job, j->d_items, j->n_items, s);
  if (st == GD_OK) st = launch_finish(j->d_job, j->d_blocks, j->n_blocks, GD_MODE_INIT, 0, s);
  if (st == GD_OK && multi && timing) st = tspan.mark_ref(s);
  int pending = -1;  // streamed dump whose copy is in flight
  for (int s1 = 0, k = 0; s1 < mc->steps_total && st == GD_OK; s1 += spd, ++k) {
    const int b = k & 1;
    if (record) {
      e = hipMemsetAsync(g.rec_n + (size_t)b * g.Rp, 0, (size_t)g.Rp * sizeof(int), s);
      if (e != hipSuccess) {
        st = hip_fail(e, "hipMemsetAsync");
        break;
      }
    }
    if (multi && (st = fork()) != GD_OK) break;
    const int s_end = std::min(mc->steps_total, s1 + spd);
    for (int step = s1; step < s_end; ++step) {
      const bool timed = timing && step % kTimedEvery == 0;
      if (multi) {
        for (int k = 0; k < np && st == GD_OK; ++k) {
          if (timed && (st = tspan.record(ps[k])) != GD_OK) break;
          if ((st = launch_pair(g.lc, j->d_job, j->d_items_part[k], j->n_items_part[k], ps[k])) != GD_OK) break;
          if (timed) st = tspan.record(ps[k]);
        }
        for (int k = 0; k < np && st == GD_OK; ++k)
          st = launch_finish(j->d_job, j->d_blocks + j->b0_part[k], j->nb_part[k], GD_MODE_STEP, step, ps[k]);
        if (st != GD_OK) break;
      } else {
        if (timed && (st = tp.begin(s)) != GD_OK) break;
        if ((st = launch_pair(g.lc, j->d_job, j->d_items, j->n_items, s)) != GD_OK) break;
        if (timed && (st = tp.end(s)) != GD_OK) break;
        if ((st = launch_finish(j->d_job, j->d_blocks, j->n_blocks, GD_MODE_STEP, step, s)) != GD_OK) break;
      }
      if (n_rounds > 0 && (step + 1) % spe == 0 && (step + 1) / spe <= n_rounds) {
        const int r = (step + 1) / spe - 1;
        if (multi && (st = join()) != GD_OK) break;
        if (tx && (st = launch_exchange(j->d_job, d_ex + r, 1, s)) != GD_OK) break;
        for (int axis = 0; axis < 2; ++axis)
          if ((axis ? px : lx) &&
              (st = launch_swap(axis, j->d_job, j->d_blocks, j->n_blocks, d_swr + r, step, nullptr,
                                nullptr, pair, s)) != GD_OK)
            break;
        if (st != GD_OK) break;
        if (multi && (st = fork()) != GD_OK) break;
      }
    }
    if (st != GD_OK) break;
    if (multi && (st = join()) != GD_OK) break;
    j->last_half = record ? b : -1;
    if (!sink) continue;
    const float *d_half = g.rec + (size_t)b * half_floats;
    const int *d_cnt = g.rec_n + (size_t)b * g.Rp;
    if (sink->blocking) {  // the reference's shape: wait, copy, hand over
      if ((e = hipStreamSynchronize(s)) != hipSuccess ||
          (e = hipMemcpy(h_half[b], d_half, half_floats * sizeof(float), hipMemcpyDeviceToHost)) !=
              hipSuccess ||
          (e = hipMemcpy(h_cnt[b], d_cnt, (size_t)g.Rp * sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess) {
        st = hip_fail(e, "record copy");
        break;
      }
      deliver(k, b);
      continue;
    }
    if ((e = hipEventRecord(ev_done[b], s)) != hipSuccess ||
        (e = hipStreamWaitEvent(j->copy_stream, ev_done[b], 0)) != hipSuccess ||
        (e = hipMemcpyAsync(h_half[b], d_half, half_floats * sizeof(float), hipMemcpyDeviceToHost,
                            j->copy_stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h_cnt[b], d_cnt, (size_t)g.Rp * sizeof(int), hipMemcpyDeviceToHost,
                            j->copy_stream)) != hipSuccess ||
        (e = hipEventRecord(ev_copied[b], j->copy_stream)) != hipSuccess) {
      st = hip_fail(e, "record stream");
      break;
    }
    if (pending >= 0) {  // the previous dump, copied while this one ran
      if ((e = hipEventSynchronize(ev_copied[pending & 1])) != hipSuccess) {
        st = hip_fail(e, "record copy");
        break;
      }
      deliver(pending, pending & 1);
    }
    pending = k;
  }
  {
    hipError_t e2 = hipStreamSynchronize(s);  // also covers steps_total == 0
    if (st == GD_OK && e2 != hipSuccess) st = hip_fail(e2, "hipStreamSynchronize");
    for (int k = 1; multi && k < np; ++k) {
      e2 = hipStreamSynchronize(ps[k]);
      if (st == GD_OK && e2 != hipSuccess) st = hip_fail(e2, "hipStreamSynchronize");
    }
    if (j->copy_stream) {
      e2 = hipStreamSynchronize(j->copy_stream);
      if (st == GD_OK && e2 != hipSuccess) st = hip_fail(e2, "record copy");
    }
  }
  if (st == GD_OK && sink && !sink->blocking && pending >= 0) deliver(pending, pending & 1);
  if (st == GD_OK && timing) st = multi ? tspan.collect(np) : tp.collect();
  const auto t1 = std::chrono::steady_clock::now();
  if (d_xt) hipFree(d_xt);
  if (st != GD_OK) return st;
  if (stats) {
    stats->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    stats->ms_pair_kernel = multi ? tspan.parts_ms : tp.ms;
    stats->ms_finish_kernel = 0.0;
    stats->n_pair_launch = multi ? tspan.n_part_launches : tp.n;
    stats->pair_evals = (double)g.R * g.L * g.P * (mc->steps_total + 1);
    stats->replica_steps = (double)g.R * mc->steps_total;
    // timed steps are 0, kTimedEvery, ...: first to last start over the steps between
    const int n_samples = multi ? tspan.n : tp.n;
    const double span_first_last = multi ? tspan.last_a - tspan.first_a : tp.first_to_last_ms;
    stats->ms_step_period =
        n_samples >= 2 ? span_first_last / ((double)(n_samples - 1) * kTimedEvery) : 0.0;
    stats->ms_pair_span = multi ? (tspan.n ? tspan.ms / tspan.n : 0.0) : (tp.n ? tp.ms / tp.n : 0.0);
  }
  return read_out(j, out);
}

gd_status gd_dock(gd_job *j, const gd_mc_param *mc, gd_replica_out *out, gd_stats *stats) {
  return dock_impl(j, mc, nullptr, out, stats);
}

gd_status gd_dock_stream(gd_job *j, const gd_mc_param *mc, gd_record_fn fn, void *user, int flags,
                         gd_replica_out *out, gd_stats *stats) {
  if (!fn || (flags & ~GD_STREAM_BLOCKING)) return GD_ERR_ARG;
  const RecSink sink = {fn, user, (flags & GD_STREAM_BLOCKING) != 0};
  return dock_impl(j, mc, &sink, out, stats);
}

gd_status gd_job_replicas(gd_job *j, gd_replica_out *out) {
  if (!j || !out) return GD_ERR_ARG;
  return gdr::read_out(j, out);
}

gd_status gd_dock_records(gd_job *j, int cap, float *rec, int *count) {
  if (!j || cap < 0 || !count || (cap > 0 && !rec)) return GD_ERR_ARG;
  const GdJob &g = j->h;
  if (!g.rec || j->last_half < 0) return GD_ERR_STATE;
  const int R = g.R, Rp = g.Rp, rc = g.rec_cap;
  const size_t half = (size_t)j->last_half;
  std::vector<int> n(Rp);
  hipError_t e = hipStreamSynchronize(j->stream);
  if (e == hipSuccess)
    e = hipMemcpy(n.data(), g.rec_n + half * Rp, (size_t)Rp * sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy rec_n");
  std::vector<float> buf((size_t)Rp * rc * GD_REC_W);
  if (!buf.empty()) {
    e = hipMemcpy(buf.data(), g.rec + half * buf.size(), buf.size() * sizeof(float),
                  hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy records");
  }
  for (int r = 0; r < R; ++r) {
    const int m = std::min(n[r], cap);
    count[r] = n[r];
    for (int k = 0; k < m; ++k)
      for (int w = 0; w < GD_REC_W; ++w)
        rec[((size_t)r * cap + k) * GD_REC_W + w] = buf[((size_t)r * rc + k) * GD_REC_W + w];
  }
  return GD_OK;
}

// Move increments as gd_dock draws them (statistical tests of the move
// generator): inc[s][r][6] for replicas r < n_rep at steps step0 + s,
// s < n_steps, computed on the device (on_device != 0, the kernels' own code)
// or on the host (the same source compiled for the host).
gd_status gd_move_draws(uint64_t seed, uint32_t job_id, int move_mode, const float move_scale[6],
                        int n_rep, uint32_t step0, int n_steps, int on_device, float *inc) {
  if (!move_scale || !inc || n_rep <= 0 || n_steps <= 0 ||
      (move_mode != GD_MOVE_REF_UNIFORM && move_mode != GD_MOVE_GAUSSIAN))
    return GD_ERR_ARG;
  const long long n = (long long)n_rep * n_steps;
  if (n > (1ll << 28)) return GD_ERR_SIZE;
  if (!on_device) {
    for (long long i = 0; i < n; ++i) {
      uint32_t w[8];
      gd_step_draws(seed, job_id, (uint32_t)(i % n_rep), step0 + (uint32_t)(i / n_rep), w);
      gd_move_incr(move_mode, move_scale, w, inc + i * 6);
    }
    return GD_OK;
  }
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  GdScale6 sc;
  for (int k = 0; k < 6; ++k) sc.v[k] = move_scale[k];
  float *d = nullptr;
  hipError_t e = hipMalloc(&d, (size_t)n * 6 * sizeof(float));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc draws");
  hipLaunchKernelGGL(gd_move_draws_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, seed,
                     job_id, move_mode, sc, n_rep, step0, (int)n, d);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(inc, d, (size_t)n * 6 * sizeof(float), hipMemcpyDeviceToHost);
  hipFree(d);
  return e == hipSuccess ? GD_OK : hip_fail(e, "gd_move_draws_kernel");
}

// Increments from given draw words: w[n][8] -> inc[n][6] (host code path).
gd_status gd_move_incr_words(int move_mode, const float move_scale[6], int n, const uint32_t *w,
                             float *inc) {
  if (!move_scale || !w || !inc || n < 0 ||
      (move_mode != GD_MOVE_REF_UNIFORM && move_mode != GD_MOVE_GAUSSIAN))
    return GD_ERR_ARG;
  for (int i = 0; i < n; ++i) gd_move_incr(move_mode, move_scale, w + (size_t)i * 8, inc + (size_t)i * 6);
  return GD_OK;
}

// Engine clock under load: every SIMD of every CU runs dependent FMAs for
// about `ms` milliseconds; the shader-clock counter's rate against the
// constant-rate counter (hipDeviceAttributeWallClockRate) is the clock.
gd_status gd_measure_sclk(int ms, double *sclk_mhz) {
  if (!sclk_mhz || ms <= 0 || ms > 2000) return GD_ERR_ARG;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  int wall_khz = 0;
  hipError_t e = hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, g_device);
  if (e != hipSuccess || wall_khz <= 0) return hip_fail(e, "hipDeviceAttributeWallClockRate");
  long long *d = nullptr;
  e = hipMalloc(&d, 2 * sizeof(long long) + sizeof(float));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc clock");
  const long long ticks = (long long)wall_khz * ms;
  hipLaunchKernelGGL(gd_clock_kernel, dim3(4 * (g_cu > 0 ? g_cu : 256)), dim3(64), 0, 0, ticks, 1.0f, d,
                     (float *)(d + 2));
  long long h[2] = {0, 0};
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
  hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "gd_clock_kernel");
  if (h[1] <= 0) return GD_ERR_STATE;
  *sclk_mhz = (double)h[0] / (double)h[1] * (double)wall_khz / 1000.0;
  return GD_OK;
}

// A plain device-to-device copy of `bytes` (a multiple of 16), `reps` times:
// the calibration for HBM traffic counters (each dispatch reads and writes
// exactly `bytes`).  *ms: mean time per copy.
gd_status gd_device_copy(size_t bytes, int reps, double *ms) {
  if (bytes == 0 || bytes % 16 != 0 || reps <= 0) return GD_ERR_ARG;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  void *a = nullptr, *b = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipMemset(a, 0, bytes);
  if (e != hipSuccess) {
    if (a) hipFree(a);
    return hip_fail(e, "gd_device_copy buffers");
  }
  const size_t n16 = bytes / 16;
  const int grid = 4 * 8 * (g_cu > 0 ? g_cu : 256);
  e = hipDeviceSynchronize();
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    hipLaunchKernelGGL(gd_copy_kernel, dim3(grid), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n16);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  const auto t1 = std::chrono::steady_clock::now();
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return hip_fail(e, "gd_copy_kernel");
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / reps;
  return GD_OK;
}

// PrintRecord-format text sink of gd_dock_stream (util_print.C:70-151).
struct gd_record_csv {
  FILE *f = nullptr;
  int lo = 0, hi = -1;
  bool failed = false;
};

gd_status gd_record_csv_open(const char *path, int rep_lo, int rep_hi, int header,
                             gd_record_csv **out) {
  if (!path || !out || rep_lo < 0 || rep_hi < rep_lo) return GD_ERR_ARG;
  *out = nullptr;
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return GD_ERR_IO;
  }
  gd_record_csv *w = new gd_record_csv();
  w->f = f;
  w->lo = rep_lo;
  w->hi = rep_hi;
  if (header) {
    static const char *names[10] = {"vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total"};
    std::fprintf(f, "rep step");
    for (int i = 0; i < 10; ++i) std::fprintf(f, " %s", names[i]);
    std::fprintf(f, "\n");
  }
  *out = w;
  return GD_OK;
}

void gd_record_csv_write(void *writer, int dump, int step_begin, int n_rep, int cap,
                         const float *rec, const int *count) {
  (void)dump;
  (void)step_begin;
  gd_record_csv *w = (gd_record_csv *)writer;
  if (!w || !w->f || !rec || !count) return;
  for (int r = w->lo; r <= w->hi && r < n_rep; ++r) {
    const int n = std::min(count[r], cap);
    for (int k = 0; k < n; ++k) {
      const float *o = rec + ((size_t)r * cap + k) * GD_REC_W;
      if (std::fprintf(w->f, "%3d %d", r, (int)o[0]) < 0) w->failed = true;
      for (int i = 0; i < 10; ++i)
        if (std::fprintf(w->f, " %.4f", o[7 + i]) < 0) w->failed = true;
      if (std::fprintf(w->f, "\n") < 0) w->failed = true;
    }
  }
}

gd_status gd_record_csv_close(gd_record_csv *w) {
  if (!w) return GD_ERR_ARG;
  bool bad = w->failed;
  if (w->f && std::fclose(w->f) != 0) bad = true;
  delete w;
  if (bad) {
    set_error("record file write failed");
    return GD_ERR_IO;
  }
  return GD_OK;
}

}  // extern "C"
