// gd_batch.hip -- many complexes in shared launches: batched docking
// (persistent work-queue pair kernel, graph-captured sweep segments), the
// tiling autotuner, and the continuous-batching serving loop.
//
// A "batch" concatenates the device structs of its jobs into one array, and
// their work items and 64-replica blocks into one list each, so one pair,
// one reduce and one finish launch advance every job by one MC step.  Each job
// keeps its own tiling, replica order, Philox stream (seed, job_id) and step
// counter (launch step - step0), so a job computes bit-identical results
// alone or inside any batch.  The reference has no counterpart: its MPI
// farm (src/main_mpi/server.C:27-66, client_gpu.C:42-72) runs one complex per
// GPU launch sequence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "gd_internal.h"

template <int LC>
__global__ void gd_pair_kernel_wq(const GdJob *, const GdItem *, int, int *);
__global__ void gd_transform_kernel(const GdJob *, const GdBlock *, int *);
__global__ void gd_reduce_kernel(const GdJob *, const GdBlock *);
__global__ void gd_finish_kernel(const GdJob *, const GdBlock *, int, int, const int *, int *);
__global__ void gd_exchange_kernel(const GdJob *, const GdExch *);

using namespace gdr;

namespace {

hipError_t launch_pair_wq(int lc, const GdJob *jobs, const GdItem *items, int n_items, int *wq,
                          int grid, hipStream_t s) {
  if (n_items <= 0) return hipSuccess;
  const dim3 g(grid), b(64 * GD_GROUP);
  switch (lc) {
    case 4:
      hipLaunchKernelGGL(gd_pair_kernel_wq<4>, g, b, 0, s, jobs, items, n_items, wq);
      break;
    case 8:
      hipLaunchKernelGGL(gd_pair_kernel_wq<8>, g, b, 0, s, jobs, items, n_items, wq);
      break;
    default:
      hipLaunchKernelGGL(gd_pair_kernel_wq<16>, g, b, 0, s, jobs, items, n_items, wq);
      break;
  }
  return hipGetLastError();
}

// Device-side lists of one batch: job structs, items, blocks, exchange table.
struct BatchLists {
  GdJob *d_jobs = nullptr;
  GdItem *d_items = nullptr;
  GdBlock *d_blocks = nullptr;
  int *d_wq = nullptr;       // work-queue counter
  int *d_base = nullptr;     // base launch step for graph replays
  GdExch *d_ex = nullptr;    // exchange entries (scratch, per launch)
  int *d_swr = nullptr;      // [n_jobs] conformer-swap round per job, -1 = none (scratch)
  std::vector<int> axes;     // [n_jobs] GD_EXCH_* bits that can act (axis bit and > 1 entries)
  int n_items = 0, n_blocks = 0, n_jobs = 0, ex_cap = 0;
  void *mem = nullptr;
  ~BatchLists() {
    if (mem) (void)hipFree(mem);
  }
  gd_status build(const std::vector<gd_job *> &jobs, const std::vector<int> &which,
                  hipStream_t s) {
    std::vector<GdJob> js;
    std::vector<GdItem> items;
    std::vector<GdBlock> blocks;
    axes.clear();
    for (int k = 0; k < (int)which.size(); ++k) {
      gd_job *j = jobs[which[k]];
      js.push_back(j->h);
      const GdJob &h = j->h;
      axes.push_back(((h.xaxes & GD_EXCH_TEMP) && h.n_tmp > 1 ? GD_EXCH_TEMP : 0) |
                     ((h.xaxes & GD_EXCH_LIG) && h.n_lig > 1 ? GD_EXCH_LIG : 0) |
                     ((h.xaxes & GD_EXCH_PRT) && h.n_prt > 1 ? GD_EXCH_PRT : 0));
      // job-local items in the job's order (chunks by decreasing KDE work, then
      // tile groups, then replica blocks)
      for (const GdItem &it : j->items) items.push_back({k, it.rb, it.a, it.b});
      for (int rb = 0; rb < j->h.Rp; rb += 64) blocks.push_back({k, rb});
    }
    // longest items first: the persistent queue then ends with small items
    std::stable_sort(items.begin(), items.end(), [&](const GdItem &a, const GdItem &b) {
      const GdJob &ja = js[a.job], &jb = js[b.job];
      return (long long)ja.ptile * ja.lc > (long long)jb.ptile * jb.lc;
    });
    n_jobs = (int)js.size();
    n_items = (int)items.size();
    n_blocks = (int)blocks.size();
    ex_cap = std::max(1, n_jobs);
    const size_t bj = (js.size() * sizeof(GdJob) + 255) / 256 * 256;
    const size_t bi = (items.size() * sizeof(GdItem) + 255) / 256 * 256;
    const size_t bb = (blocks.size() * sizeof(GdBlock) + 255) / 256 * 256;
    const size_t be = ((size_t)ex_cap * sizeof(GdExch) + 255) / 256 * 256;
    const size_t bs = ((size_t)ex_cap * sizeof(int) + 255) / 256 * 256;
    if (mem) (void)hipFree(mem);
    mem = nullptr;
    hipError_t e = hipMalloc(&mem, bj + bi + bb + be + bs + 512);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc batch");
    char *p = (char *)mem;
    d_jobs = (GdJob *)p;
    d_items = (GdItem *)(p + bj);
    d_blocks = (GdBlock *)(p + bj + bi);
    d_ex = (GdExch *)(p + bj + bi + bb);
    d_swr = (int *)(p + bj + bi + bb + be);
    d_wq = (int *)(p + bj + bi + bb + be + bs);
    d_base = d_wq + 64;
    if ((e = hipMemcpyAsync(d_jobs, js.data(), js.size() * sizeof(GdJob), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(GdItem), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(d_blocks, blocks.data(), blocks.size() * sizeof(GdBlock), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemsetAsync(d_wq, 0, 512, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return hip_fail(e, "batch upload");
    return GD_OK;
  }
};

int pair_grid(int waves_per_cu) {  // work groups of GD_GROUP waves
  const int cu = std::max(1, n_cu());
  return std::max(1, cu * std::max(GD_GROUP, waves_per_cu) / GD_GROUP);
}

// Jobs in shared launches (batches, serving) run 4 atoms per lane and the
// table's tile size for their size bucket (shared_ptile): a function of the
// complex only, so a job's bits do not depend on what it runs with.
constexpr int kSharedLc = 4;

// Bring a job to its shared-launch tiling (ptile_override > 0: that tile size
// instead, for the sweeps of gd_autotune_shared).
gd_status shared_tiling(gd_job *j, int ptile_override) {
  const int pt = ptile_override > 0 ? std::min(j->h.P, ptile_override) : shared_ptile(j->h.L, j->h.P);
  if (j->h.lc == kSharedLc && j->h.ptile == pt) return GD_OK;
  gd_status st = apply_tiling(j, kSharedLc, pt);
  if (st != GD_OK) return st;
  return sync_job_struct(j);
}

// Work of one exchange point: temperature-exchange entries and, per job of the
// lists, the round of its conformer swaps (-1 = none).
struct ExchPoint {
  std::vector<GdExch> ex;
  std::vector<int> swr;
  int axes = 0;  // swap axes to launch (GD_EXCH_LIG / GD_EXCH_PRT)
  bool empty() const { return ex.empty() && axes == 0; }
  // round[q] >= 0 marks list job q as being at an exchange point
  void set(const BatchLists &B, const std::vector<int> &round) {
    ex.clear();
    swr.assign(B.n_jobs, -1);
    axes = 0;
    for (int q = 0; q < B.n_jobs; ++q) {
      if (round[q] < 0) continue;
      if (B.axes[q] & GD_EXCH_TEMP) ex.push_back({q, round[q]});
      if (B.axes[q] & (GD_EXCH_LIG | GD_EXCH_PRT)) {
        swr[q] = round[q];
        axes |= B.axes[q] & (GD_EXCH_LIG | GD_EXCH_PRT);
      }
    }
  }
};

// Exchange kernels of one exchange point, tables already on the device:
// temperature exchange (n_ex entries of B.d_ex), then the swap sub-steps.
gd_status exchange_launch(const BatchLists &B, int lc, int grid, int n_ex, int axes, int launch_step,
                          const int *base, hipStream_t s) {
  if (n_ex > 0) {
    hipLaunchKernelGGL(gd_exchange_kernel, dim3(n_ex), dim3(256), 0, s, B.d_jobs, B.d_ex);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "exchange kernel launch");
  }
  const std::function<gd_status()> pair = [&]() {
    hipError_t e = launch_pair_wq(lc, B.d_jobs, B.d_items, B.n_items, B.d_wq, grid, s);
    return e == hipSuccess ? GD_OK : hip_fail(e, "pair kernel (queue) launch");
  };
  for (int axis = 0; axis < 2; ++axis)
    if (axes & (GD_EXCH_LIG << axis)) {
      gd_status st = launch_swap(axis, B.d_jobs, B.d_blocks, B.n_blocks, B.d_swr, launch_step, base,
                                 B.d_wq, pair, s);
      if (st != GD_OK) return st;
    }
  return GD_OK;
}

// Upload the tables of an exchange point (the caller synchronises before the
// tables are rewritten).
gd_status exchange_upload(const BatchLists &B, const ExchPoint &xp, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (!xp.ex.empty())
    e = hipMemcpyAsync(B.d_ex, xp.ex.data(), xp.ex.size() * sizeof(GdExch), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && xp.axes)
    e = hipMemcpyAsync(B.d_swr, xp.swr.data(), xp.swr.size() * sizeof(int), hipMemcpyHostToDevice, s);
  return e == hipSuccess ? GD_OK : hip_fail(e, "hipMemcpyAsync exchange tables");
}

// One MC step of every job in the lists (pair + finish), optional exchange point.
gd_status batch_step(const BatchLists &B, int lc, int grid, int launch_step, const int *base,
                     const ExchPoint *xp, hipStream_t s, KTimer *tp) {
  const bool timed = tp && launch_step % kTimedEvery == 0;
  if (timed && tp->begin(s) != GD_OK) return GD_ERR_HIP;
  hipError_t e = launch_pair_wq(lc, B.d_jobs, B.d_items, B.n_items, B.d_wq, grid, s);
  if (e != hipSuccess) return hip_fail(e, "pair kernel (queue) launch");
  if (timed && tp->end(s) != GD_OK) return GD_ERR_HIP;
  hipLaunchKernelGGL(gd_reduce_kernel, dim3(B.n_blocks * GD_RK), dim3(256), 0, s, B.d_jobs, B.d_blocks);
  hipLaunchKernelGGL(gd_finish_kernel, dim3(B.n_blocks), dim3(1024), 0, s, B.d_jobs, B.d_blocks,
                     GD_MODE_STEP, launch_step, base, B.d_wq);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "finish kernel launch");
  if (xp && !xp->empty()) {
    gd_status st = exchange_upload(B, *xp, s);
    if (st != GD_OK) return st;
    return exchange_launch(B, lc, grid, (int)xp->ex.size(), xp->axes, launch_step, base, s);
  }
  return GD_OK;
}

gd_status batch_init(const BatchLists &B, int lc, int grid, hipStream_t s) {
  hipLaunchKernelGGL(gd_transform_kernel, dim3(B.n_blocks), dim3(256), 0, s, B.d_jobs, B.d_blocks,
                     B.d_wq);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "transform kernel launch");
  if ((e = launch_pair_wq(lc, B.d_jobs, B.d_items, B.n_items, B.d_wq, grid, s)) != hipSuccess)
    return hip_fail(e, "pair kernel (queue) launch");
  hipLaunchKernelGGL(gd_reduce_kernel, dim3(B.n_blocks * GD_RK), dim3(256), 0, s, B.d_jobs, B.d_blocks);
  hipLaunchKernelGGL(gd_finish_kernel, dim3(B.n_blocks), dim3(1024), 0, s, B.d_jobs, B.d_blocks,
                     GD_MODE_INIT, 0, (const int *)nullptr, B.d_wq);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "finish kernel launch");
  return GD_OK;
}

// Batched docking (gd_batch_dock); the overrides (> 0) replace the shared
// tile size and the persistent grid's waves per CU (autotuner sweeps).
gd_status batch_dock_impl(gd_job **jobs_in, int n_jobs, const gd_mc_param *mc, gd_replica_out *out,
                          gd_stats *stats, int ptile_override, int waves_override) {
  if (!jobs_in || n_jobs <= 0 || !mc) return GD_ERR_ARG;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  std::vector<gd_job *> jobs(jobs_in, jobs_in + n_jobs);
  for (gd_job *j : jobs)
    if (!j) return GD_ERR_ARG;
  const int lc = kSharedLc;
  for (gd_job *j : jobs)
    if ((st = shared_tiling(j, ptile_override)) != GD_OK) return st;
  hipStream_t s = jobs[0]->stream;
  int steps_max = 0;
  for (int k = 0; k < n_jobs; ++k) {
    if (mc[k].steps_total < 0 || mc[k].steps_per_exchange < 0) return GD_ERR_ARG;
    set_mc(jobs[k]->h, &mc[k]);
    jobs[k]->h.rec = nullptr;
    jobs[k]->h.rec_cap = 0;
    jobs[k]->h.record = 0;
    if ((st = reset_replicas(jobs[k])) != GD_OK) return st;
    steps_max = std::max(steps_max, mc[k].steps_total);
  }
  // jobs run from launch step 0; a job whose steps are done leaves the lists
  std::vector<int> active(n_jobs);
  for (int k = 0; k < n_jobs; ++k) active[k] = k;
  BatchLists B;
  if ((st = B.build(jobs, active, s)) != GD_OK) return st;
  const int grid = pair_grid(waves_override > 0 ? waves_override : pair_waves_per_cu(lc));
  KTimer tp;
  const auto t0 = std::chrono::steady_clock::now();
  if ((st = batch_init(B, lc, grid, s)) != GD_OK) return st;

  // Segments of identical shape are captured once into a HIP graph and
  // replayed with the base step in device memory (no memset nodes).
  int seg = 0;
  bool uniform = true;
  for (int k = 0; k < n_jobs; ++k)
    uniform = uniform && mc[k].steps_total == mc[0].steps_total &&
              mc[k].steps_per_exchange == mc[0].steps_per_exchange;
  if (uniform) seg = mc[0].steps_per_exchange > 0 ? mc[0].steps_per_exchange : 50;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  double pair_evals = 0.0, rep_steps = 0.0;
  for (int k = 0; k < n_jobs; ++k) {
    const GdJob &g = jobs[k]->h;
    pair_evals += (double)g.R * g.L * g.P * (mc[k].steps_total + 1);
    rep_steps += (double)g.R * mc[k].steps_total;
  }
  int step = 0;
  ExchPoint xp;
  std::vector<int> round;
  while (step < steps_max && st == GD_OK) {
    // graph path: whole segment, every job active, exchange point at its end
    if (uniform && seg > 1 && step + seg <= steps_max && !stats) {
      hipError_t e = hipMemcpyAsync(B.d_base, &step, sizeof(int), hipMemcpyHostToDevice, s);
      if (e != hipSuccess) {
        st = hip_fail(e, "hipMemcpyAsync base");
        break;
      }
      // the segment's exchange tables (the round is (step + seg) / seg - 1 for
      // every job), uploaded before each replay
      round.assign(B.n_jobs, mc[0].steps_per_exchange > 0 ? (step + seg) / seg - 1 : -1);
      xp.set(B, round);
      if (!gexec) {
        if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
          st = hip_fail(e, "hipStreamBeginCapture");
          break;
        }
        for (int k = 0; k < seg && st == GD_OK; ++k) {
          st = batch_step(B, lc, grid, k, B.d_base, nullptr, s, nullptr);
          if (st == GD_OK && k == seg - 1 && !xp.empty())
            st = exchange_launch(B, lc, grid, (int)xp.ex.size(), xp.axes, k, B.d_base, s);
        }
        hipError_t e2 = hipStreamEndCapture(s, &graph);
        if (st == GD_OK && e2 != hipSuccess) st = hip_fail(e2, "hipStreamEndCapture");
        if (st != GD_OK) break;
        if ((e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0)) != hipSuccess) {
          st = hip_fail(e, "hipGraphInstantiate");
          break;
        }
      }
      if (!xp.empty() && (st = exchange_upload(B, xp, s)) != GD_OK) break;
      if ((e = hipGraphLaunch(gexec, s)) != hipSuccess) {
        st = hip_fail(e, "hipGraphLaunch");
        break;
      }
      // the host-side copies into d_base and the exchange tables must not be
      // overwritten while the graph may still read them: keep one segment in flight
      if ((e = hipStreamSynchronize(s)) != hipSuccess) {
        st = hip_fail(e, "hipStreamSynchronize");
        break;
      }
      step += seg;
      continue;
    }
    // eager path (timed runs, ragged tails, jobs with different lengths)
    std::vector<int> act;
    for (int k = 0; k < n_jobs; ++k)
      if (step < mc[k].steps_total) act.push_back(k);
    if (act.size() != (size_t)B.n_jobs) {
      // shrink the lists to the jobs still running; their step0 stays 0
      if ((st = B.build(jobs, act, s)) != GD_OK) break;
    }
    round.assign(act.size(), -1);
    for (int q = 0; q < (int)act.size(); ++q) {
      const int spe = mc[act[q]].steps_per_exchange;
      if (spe > 0 && (step + 1) % spe == 0) round[q] = (step + 1) / spe - 1;
    }
    xp.set(B, round);
    st = batch_step(B, lc, grid, step, nullptr, &xp, s, stats ? &tp : nullptr);
    if (st != GD_OK) break;
    if (!xp.empty() || (stats && (step + 1) % 64 == 0)) {
      hipError_t e = hipStreamSynchronize(s);  // exchange tables are rewritten next time
      if (e != hipSuccess) {
        st = hip_fail(e, "hipStreamSynchronize");
        break;
      }
      if (stats && (st = tp.collect()) != GD_OK) break;
    }
    ++step;
  }
  hipError_t e = hipStreamSynchronize(s);
  if (st == GD_OK && e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize");
  if (st == GD_OK && stats) st = tp.collect();
  const auto t1 = std::chrono::steady_clock::now();
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (st != GD_OK) return st;
  if (stats) {
    stats->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    stats->ms_pair_kernel = tp.ms;
    stats->ms_finish_kernel = 0.0;
    stats->n_pair_launch = tp.n;
    stats->pair_evals = pair_evals;
    stats->replica_steps = rep_steps;
    // the batch's timed runs synchronise every 64 steps: the period on the host clock
    stats->ms_step_period = steps_max > 0 ? stats->ms_total / steps_max : 0.0;
    stats->ms_pair_span = tp.n ? tp.ms / tp.n : 0.0;
  }
  if (out)
    for (int k = 0; k < n_jobs; ++k)
      if ((st = read_out(jobs[k], &out[k])) != GD_OK) return st;
  return GD_OK;
}

}  // namespace

extern "C" gd_status gd_batch_dock(gd_job **jobs, int n_jobs, const gd_mc_param *mc,
                                   gd_replica_out *out, gd_stats *stats) {
  return batch_dock_impl(jobs, n_jobs, mc, out, stats, 0, 0);
}

// ---------------------------------------------------------------------------
// Autotuner (the tiling table of gd_tiling_table.h is written from these sweeps
// by geauxdock_amd/tune.py).  The reference's only counterpart is the disabled
// CalcThreadBlockShape (backend_gpu/kernel_cuda_l2_util.cu:40-67), which picks
// a block shape by a formula.  Neither sweep touches process-wide state: a
// job's tiling elsewhere stays a function of the job.
namespace {
gd_mc_param tune_mc(int steps) {
  gd_mc_param mc;
  std::memset(&mc, 0, sizeof mc);
  mc.steps_total = steps;
  mc.steps_per_dump = steps;
  mc.move_mode = GD_MOVE_GAUSSIAN;
  for (int i = 0; i < 6; ++i) mc.move_scale[i] = i < 3 ? 0.2f : 0.05f;
  mc.seed = 1;
  return mc;
}
}  // namespace

// One job per launch: atoms per lane {4, 8} x protein tile counts, each run
// for `steps` MC steps on the job's streams; the shortest steady-state MC
// step period (HIP events at every 16th step: 33 steps or more; fewer fall
// back to wall time per step) wins and is applied to this job (only to it).
extern "C" gd_status gd_autotune(gd_job *j, int steps, gd_tiling *best, double *best_ms) {
  if (!j || steps <= 0) return GD_ERR_ARG;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  const GdJob g0 = j->h;
  const int P = g0.P;
  std::vector<std::pair<int, int>> cand;  // (lc, ptile)
  for (int lc : {4, 8})
    for (int nP : {16, 24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 112, 128, 160}) {
      const int pt = (P + nP - 1) / nP;
      if (pt >= 4 && std::find(cand.begin(), cand.end(), std::make_pair(lc, pt)) == cand.end())
        cand.push_back({lc, pt});
    }
  {
    int dlc, dpt;
    default_tiling(g0.L, P, g0.R, dlc, dpt);
    if (std::find(cand.begin(), cand.end(), std::make_pair(dlc, dpt)) == cand.end())
      cand.push_back({dlc, dpt});
  }
  const gd_mc_param mc = tune_mc(steps);
  double best_t = 1e300;
  std::pair<int, int> bestc = {g0.lc, g0.ptile};
  for (auto &c : cand) {
    if ((st = apply_tiling(j, c.first, c.second)) != GD_OK) return st;
    if ((st = sync_job_struct(j)) != GD_OK) return st;
    gd_stats s;
    if ((st = gd_dock(j, &mc, nullptr, &s)) != GD_OK) return st;
    // what the job pays per MC step: the steady-state step period on the
    // job's streams (on several streams a pair launch is only part of a step)
    const double t = s.ms_step_period > 0.0 ? s.ms_step_period : s.ms_total / steps;
    if (t < best_t) {
      best_t = t;
      bestc = c;
    }
  }
  if ((st = apply_tiling(j, bestc.first, bestc.second)) != GD_OK) return st;
  if ((st = sync_job_struct(j)) != GD_OK) return st;
  if (best) {
    best->atoms_per_lane = bestc.first;
    best->tile_points = bestc.second;
    best->waves_per_cu = 0;
  }
  if (best_ms) *best_ms = best_t;
  return GD_OK;
}

// Shared launches: protein tile points x persistent-grid waves per CU (4 atoms
// per lane), each candidate a whole batched run of `steps` MC steps of the
// given jobs (graph-captured as gd_batch_dock runs them); the shortest run
// wins.  The jobs are left at the shared tiling of the table.
extern "C" gd_status gd_autotune_shared(gd_job **jobs, int n_jobs, int steps, gd_tiling *best,
                                        double *best_ms) {
  if (!jobs || n_jobs <= 0 || steps <= 0) return GD_ERR_ARG;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  std::vector<gd_mc_param> mc(n_jobs, tune_mc(steps));
  for (int k = 0; k < n_jobs; ++k) mc[k].job_id = (uint32_t)k;
  double best_t = 1e300;
  int best_pt = 0, best_w = 0;
  for (int pt : {32, 40, 48, 64, 80, 96, 128})
    for (int w : {16, 24, 32}) {
      if ((st = batch_dock_impl(jobs, n_jobs, mc.data(), nullptr, nullptr, pt, w)) != GD_OK) return st;
      const auto t0 = std::chrono::steady_clock::now();
      if ((st = batch_dock_impl(jobs, n_jobs, mc.data(), nullptr, nullptr, pt, w)) != GD_OK) return st;
      const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (t < best_t) {
        best_t = t;
        best_pt = pt;
        best_w = w;
      }
    }
  for (int k = 0; k < n_jobs; ++k)
    if ((st = shared_tiling(jobs[k], 0)) != GD_OK) return st;
  if (best) {
    best->atoms_per_lane = kSharedLc;
    best->tile_points = best_pt;
    best->waves_per_cu = best_w;
  }
  if (best_ms) *best_ms = best_t;
  return GD_OK;
}

// The tiling the library gives a job of these sizes: one job per launch
// (shared == 0) or inside shared launches.
extern "C" gd_status gd_tiling_lookup(int L, int P, int R, int shared, gd_tiling *t) {
  if (!t || L < 1 || P < 1 || R < 1) return GD_ERR_ARG;
  if (shared) {
    t->atoms_per_lane = kSharedLc;
    t->tile_points = shared_ptile(L, P);
    t->waves_per_cu = pair_waves_per_cu(kSharedLc);
  } else {
    int lc, pt;
    default_tiling(L, P, R, lc, pt);
    t->atoms_per_lane = lc;
    t->tile_points = pt;
    t->waves_per_cu = 0;
  }
  return GD_OK;
}

// ---------------------------------------------------------------------------
// Serving loop: requests queue on the host, are admitted when their device
// footprint fits the HBM budget and an active slot is free, run in shared
// launches with every other active job (continuous batching), and retire
// as soon as their own step count is reached.
struct ServJob {
  int64_t ticket;
  gd_job_desc desc;
  gd_mc_param mc;
  gd_replica_out *out;
  size_t foot;
  gd_job *job = nullptr;
  gd_host_job *host = nullptr;  // file requests: the parsed complex behind desc
  int done_steps = 0;
  int step0 = 0;
  double t_submit = 0, t_admit = 0;
  bool init_done = false;
  bool converged = false;
  std::vector<float> best_hist;  // lowest accepted energy at each sweep's end
  ~ServJob() {
    if (host) gd_host_job_free(host);
  }
};

// A request given by its files, waiting for the loader thread.
struct FileReq {
  int64_t ticket;
  std::string sdf, ff, pdb, id;
  const gd_enepara *ene;
  int n_tmp;
  float floor_temp, ceiling_temp;
  gd_mc_param mc;
  double t_submit;
};

// Server-owned final replica state of a file request (until gd_server_take).
struct HeldOut {
  int n_rep = 0;
  bool done = false;
  std::vector<float> mm, en, best;
  std::vector<int> tmp, nacc, nexc, acc, nsl, nsp;
  gd_replica_out view{};
  explicit HeldOut(int n)
      : n_rep(n), mm((size_t)n * 6), en((size_t)n * 10), best(n), tmp(n), nacc(n), nexc(n), acc(n),
        nsl(n), nsp(n) {
    view.movematrix = mm.data();
    view.energy = en.data();
    view.best_energy = best.data();
    view.idx_tmp = tmp.data();
    view.n_accept = nacc.data();
    view.n_exchange = nexc.data();
    view.is_accept = acc.data();
    view.n_swap_lig = nsl.data();
    view.n_swap_prt = nsp.data();
  }
};

struct gd_server {
  gd_server_config cfg;
  // mu: pending, results, used, next_ticket, held, cancelled, loading
  std::mutex mu;
  std::deque<ServJob *> pending;
  std::vector<gd_job_result> results;
  size_t used = 0;
  int64_t next_ticket = 1;
  std::map<int64_t, HeldOut *> held;
  std::set<int64_t> cancelled;   // file requests cancelled while the loader parsed them
  int64_t loading = 0;           // ticket the loader is parsing (0: none)
  // smu: serialises gd_server_step with cancellations of active jobs, and
  // guards active, lists, dirty, launch_step (never held by submit or results)
  std::mutex smu;
  std::vector<ServJob *> active;
  int launch_step = 0;
  hipStream_t stream = nullptr;
  BatchLists lists;
  bool dirty = true;
  int lc = kSharedLc;
  std::chrono::steady_clock::time_point t0;
  // host input pipeline: file requests parsed by a loader thread
  std::mutex fmu;                // files, stop
  std::condition_variable fcv;
  std::deque<FileReq *> files;
  bool stop = false;
  std::thread loader;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace gdr {
size_t footprint(const gd_job_desc *d, bool shared);
}

// Loader thread: parses file requests outside every lock but its queue's, so
// parsing overlaps the device work of gd_server_step, then queues the parsed
// job for admission (or reports it cancelled, if it was while parsing).
static void server_loader(gd_server *s) {
  for (;;) {
    FileReq *r;
    {
      std::unique_lock<std::mutex> lk(s->fmu);
      s->fcv.wait(lk, [&] { return s->stop || !s->files.empty(); });
      if (s->stop) return;
      r = s->files.front();
      s->files.pop_front();
      std::lock_guard<std::mutex> lm(s->mu);
      s->loading = r->ticket;
    }
    gd_host_job *h = nullptr;
    gd_job_desc d;
    std::memset(&d, 0, sizeof d);
    size_t foot = 0;
    gd_status st = gd_load_complex(r->sdf.c_str(), r->ff.c_str(), r->pdb.c_str(), r->id.c_str(),
                                   r->ene, r->n_tmp, r->floor_temp, r->ceiling_temp, &h);
    if (st == GD_OK) st = gd_host_job_desc(h, &d);
    if (st == GD_OK) {
      foot = footprint(&d, true);
      if (foot == 0) st = GD_ERR_ARG;
      else if (foot > s->cfg.hbm_budget) st = GD_ERR_BUSY;  // can never be admitted
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->loading = 0;
    if (s->cancelled.erase(r->ticket)) st = GD_ERR_STATE;  // cancelled while parsed
    if (st != GD_OK) {
      if (h) gd_host_job_free(h);
      s->results.push_back({r->ticket, st, r->t_submit, 0.0, s->now(), 0.0});
      delete r;
      continue;
    }
    HeldOut *o = new HeldOut(d.n_lig * d.n_prt * d.n_tmp);
    s->held[r->ticket] = o;
    ServJob *j = new ServJob();
    j->ticket = r->ticket;
    j->desc = d;
    j->mc = r->mc;
    j->out = &o->view;
    j->foot = foot;
    j->host = h;
    j->t_submit = r->t_submit;
    s->pending.push_back(j);
    delete r;
  }
}


extern "C" gd_status gd_server_create(const gd_server_config *cfg, gd_server **out) {
  if (!cfg || !out || cfg->sweep_steps <= 0 || cfg->max_active <= 0 || cfg->converge_sweeps < 0 ||
      !(cfg->converge_tol >= 0.0f))
    return GD_ERR_ARG;
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  gd_server *s = new gd_server();
  s->cfg = *cfg;
  s->t0 = std::chrono::steady_clock::now();
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = s;
  return GD_OK;
}

extern "C" gd_status gd_server_destroy(gd_server *s) {
  if (!s) return GD_OK;
  {
    std::lock_guard<std::mutex> lk(s->fmu);
    s->stop = true;
  }
  s->fcv.notify_all();
  if (s->loader.joinable()) s->loader.join();
  for (FileReq *r : s->files) delete r;
  for (auto &kv : s->held) delete kv.second;
  (void)hipStreamSynchronize(s->stream);
  for (ServJob *j : s->active) {
    gd_job_destroy(j->job);
    delete j;
  }
  for (ServJob *j : s->pending) delete j;
  (void)hipStreamDestroy(s->stream);
  delete s;
  return GD_OK;
}

extern "C" double gd_server_clock(gd_server *s) { return s ? s->now() : 0.0; }

extern "C" gd_status gd_server_submit(gd_server *s, const gd_job_desc *desc, const gd_mc_param *mc,
                                      gd_replica_out *out, int64_t *ticket) {
  if (!s || !desc || !mc || mc->steps_total < 0) return GD_ERR_ARG;
  const size_t foot = footprint(desc, true);
  if (foot == 0) return GD_ERR_ARG;
  if (foot > s->cfg.hbm_budget) return GD_ERR_BUSY;  // can never be admitted
  ServJob *j = new ServJob();
  j->desc = *desc;
  j->mc = *mc;
  j->out = out;
  j->foot = foot;
  std::lock_guard<std::mutex> lk(s->mu);  // never held across a device sweep
  j->ticket = s->next_ticket++;
  j->t_submit = s->now();
  s->pending.push_back(j);
  if (ticket) *ticket = j->ticket;
  return GD_OK;
}

extern "C" gd_status gd_server_submit_files(gd_server *s, const gd_request_files *req,
                                            const gd_mc_param *mc, int64_t *ticket) {
  if (!s || !req || !mc || mc->steps_total < 0 || !req->sdf || !req->ff || !req->pdb ||
      !req->ligand_id || !req->ene || req->n_tmp < 1)
    return GD_ERR_ARG;
  FileReq *r = new FileReq();
  r->sdf = req->sdf;
  r->ff = req->ff;
  r->pdb = req->pdb;
  r->id = req->ligand_id;
  r->ene = req->ene;
  r->n_tmp = req->n_tmp;
  r->floor_temp = req->floor_temp;
  r->ceiling_temp = req->ceiling_temp;
  r->mc = *mc;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    r->t_submit = s->now();
  }
  const int64_t t = r->ticket;
  {
    std::lock_guard<std::mutex> lk(s->fmu);
    s->files.push_back(r);
    if (!s->loader.joinable()) s->loader = std::thread(server_loader, s);
  }
  s->fcv.notify_one();
  if (ticket) *ticket = t;
  return GD_OK;
}

extern "C" gd_status gd_server_take(gd_server *s, int64_t ticket, int *n_rep, gd_replica_out *out) {
  if (!s) return GD_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->held.find(ticket);
  if (it == s->held.end()) return GD_ERR_ARG;
  HeldOut *o = it->second;
  if (!o->done) return GD_ERR_STATE;
  if (n_rep) *n_rep = o->n_rep;
  if (!out) return GD_OK;
  const size_t n = (size_t)o->n_rep;
  auto cp = [](void *dst, const void *src, size_t bytes) {
    if (dst) std::memcpy(dst, src, bytes);
  };
  cp(out->movematrix, o->mm.data(), n * 6 * sizeof(float));
  cp(out->energy, o->en.data(), n * 10 * sizeof(float));
  cp(out->best_energy, o->best.data(), n * sizeof(float));
  cp(out->idx_tmp, o->tmp.data(), n * sizeof(int));
  cp(out->n_accept, o->nacc.data(), n * sizeof(int));
  cp(out->n_exchange, o->nexc.data(), n * sizeof(int));
  cp(out->is_accept, o->acc.data(), n * sizeof(int));
  cp(out->n_swap_lig, o->nsl.data(), n * sizeof(int));
  cp(out->n_swap_prt, o->nsp.data(), n * sizeof(int));
  delete o;
  s->held.erase(it);
  return GD_OK;
}

namespace {
// A request that will not run: its result is the status, and a file request
// leaves nothing to take.  Caller holds mu.
void fail_request(gd_server *s, int64_t ticket, gd_status st, double t_submit, double t_admit) {
  s->results.push_back({ticket, st, t_submit, t_admit, s->now(), 0.0});
  auto h = s->held.find(ticket);
  if (h != s->held.end()) {
    delete h->second;
    s->held.erase(h);
  }
}
}  // namespace

extern "C" gd_status gd_server_cancel(gd_server *s, int64_t ticket) {
  if (!s) return GD_ERR_ARG;
  {  // a file request not parsed yet, or being parsed
    std::lock_guard<std::mutex> lf(s->fmu);
    for (auto it = s->files.begin(); it != s->files.end(); ++it)
      if ((*it)->ticket == ticket) {
        FileReq *r = *it;
        s->files.erase(it);
        std::lock_guard<std::mutex> lk(s->mu);
        s->results.push_back({ticket, GD_ERR_STATE, r->t_submit, 0.0, s->now(), 0.0});
        delete r;
        return GD_OK;
      }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->loading == ticket) {  // the loader reports it cancelled when done parsing
      s->cancelled.insert(ticket);
      return GD_OK;
    }
  }
  {  // a pending request
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
      if ((*it)->ticket == ticket) {
        ServJob *j = *it;
        s->pending.erase(it);
        fail_request(s, ticket, GD_ERR_STATE, j->t_submit, 0.0);
        delete j;
        return GD_OK;
      }
  }
  // an active job: wait for a running sweep
  std::lock_guard<std::mutex> ls(s->smu);
  for (auto it = s->active.begin(); it != s->active.end(); ++it)
    if ((*it)->ticket == ticket) {
      (void)hipStreamSynchronize(s->stream);
      ServJob *j = *it;
      s->active.erase(it);
      s->dirty = true;
      std::lock_guard<std::mutex> lk(s->mu);
      fail_request(s, ticket, GD_ERR_STATE, j->t_submit, j->t_admit);
      s->results.back().replica_steps = (double)j->job->h.R * j->done_steps;
      s->used -= j->foot;
      gd_job_destroy(j->job);
      delete j;
      return GD_OK;
    }
  return GD_ERR_ARG;
}

extern "C" gd_status gd_server_usage(gd_server *s, size_t *used, int *n_pending) {
  if (!s) return GD_ERR_ARG;
  int nf;
  {
    std::lock_guard<std::mutex> lk(s->fmu);
    nf = (int)s->files.size();
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (used) *used = s->used;
  if (n_pending) *n_pending = (int)s->pending.size() + nf + (s->loading ? 1 : 0);
  return GD_OK;
}

extern "C" gd_status gd_server_results(gd_server *s, gd_job_result *res, int cap, int *n) {
  if (!s || !n || (cap > 0 && !res)) return GD_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  const int m = std::min(cap, (int)s->results.size());
  for (int i = 0; i < m; ++i) res[i] = s->results[i];
  s->results.erase(s->results.begin(), s->results.begin() + m);
  *n = m;
  return GD_OK;
}

namespace {
// Lowest energy any replica of the job has accepted (device best[] array).
gd_status job_best(gd_job *j, hipStream_t s, float *out) {
  std::vector<float> b(j->h.R);
  hipError_t e = hipMemcpyAsync(b.data(), j->h.best, b.size() * sizeof(float), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy best");
  float m = b.empty() ? 0.0f : b[0];
  for (float v : b) m = std::min(m, v);
  *out = m;
  return GD_OK;
}
}  // namespace

extern "C" gd_status gd_server_step(gd_server *s, int *n_active, int *n_done) {
  if (!s) return GD_ERR_ARG;
  std::lock_guard<std::mutex> ls(s->smu);
  gd_status st = ensure_device();
  if (st != GD_OK) return st;
  hipStream_t hs = s->stream;
  // (1) admission: FIFO while the estimated footprint fits and a slot is free;
  // an admitted job is then charged its actual device bytes
  std::vector<ServJob *> admitted;
  for (;;) {
    ServJob *j = nullptr;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->pending.empty() || (int)s->active.size() + (int)admitted.size() >= s->cfg.max_active ||
          s->used + s->pending.front()->foot > s->cfg.hbm_budget)
        break;
      j = s->pending.front();
      s->pending.pop_front();
      s->used += j->foot;  // held while the job is created
    }
    st = gd_job_create(&j->desc, &j->job);
    if (st == GD_OK) st = shared_tiling(j->job, 0);  // serving tiling: a function of the complex only
    if (st == GD_OK) {
      set_mc(j->job->h, &j->mc);
      j->job->h.record = 0;
      j->job->h.step0 = s->launch_step;
      st = reset_replicas(j->job);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->used -= j->foot;
    if (st != GD_OK) {
      fail_request(s, j->ticket, st, j->t_submit, s->now());
      if (j->job) gd_job_destroy(j->job);
      delete j;
      continue;
    }
    j->foot = gd_job_device_bytes(j->job);
    s->used += j->foot;
    j->t_admit = s->now();
    j->step0 = s->launch_step;
    admitted.push_back(j);
  }
  // (2) initial-energy pass of the newly admitted jobs, in one batch
  if (!admitted.empty()) {
    std::vector<gd_job *> nj;
    std::vector<int> idx;
    for (ServJob *j : admitted) {
      idx.push_back((int)nj.size());
      nj.push_back(j->job);
    }
    BatchLists B;
    st = B.build(nj, idx, hs);
    if (st == GD_OK) st = batch_init(B, s->lc, pair_grid(pair_waves_per_cu(s->lc)), hs);
    if (st == GD_OK) {
      hipError_t e = hipStreamSynchronize(hs);
      if (e != hipSuccess) st = hip_fail(e, "init pass");
    }
    if (st != GD_OK) {  // none of them runs: report each
      std::lock_guard<std::mutex> lk(s->mu);
      for (ServJob *j : admitted) {
        fail_request(s, j->ticket, st, j->t_submit, j->t_admit);
        s->used -= j->foot;
        gd_job_destroy(j->job);
        delete j;
      }
      return st;
    }
    for (ServJob *j : admitted) s->active.push_back(j);
    s->dirty = true;
  }
  // (3) advance every active job by up to sweep_steps (no lock held but smu:
  // submits and result queries proceed meanwhile)
  for (int k = 0; k < s->cfg.sweep_steps && st == GD_OK; ++k) {
    std::vector<int> act;
    std::vector<gd_job *> jobs;
    for (ServJob *j : s->active) jobs.push_back(j->job);
    for (int q = 0; q < (int)s->active.size(); ++q)
      if (s->active[q]->done_steps < s->active[q]->mc.steps_total && !s->active[q]->converged)
        act.push_back(q);
    if (act.empty()) break;
    if (s->dirty || (int)act.size() != s->lists.n_jobs) {
      for (int q : act) jobs[q]->h.step0 = s->active[q]->step0;
      if ((st = s->lists.build(jobs, act, hs)) != GD_OK) return st;
      s->dirty = false;
    }
    std::vector<int> round(act.size(), -1);
    for (int q = 0; q < (int)act.size(); ++q) {
      ServJob *j = s->active[act[q]];
      const int spe = j->mc.steps_per_exchange;
      if (spe > 0 && (j->done_steps + 1) % spe == 0) round[q] = (j->done_steps + 1) / spe - 1;
    }
    ExchPoint xp;
    xp.set(s->lists, round);
    st = batch_step(s->lists, s->lc, pair_grid(pair_waves_per_cu(s->lc)), s->launch_step, nullptr, &xp, hs, nullptr);
    if (!xp.empty()) (void)hipStreamSynchronize(hs);  // exchange tables are reused next step
    ++s->launch_step;
    for (int q : act) s->active[q]->done_steps += 1;
    // jobs whose step counts end here leave the lists at the next step
    bool any_done = false;
    for (int q : act)
      any_done = any_done || s->active[q]->done_steps >= s->active[q]->mc.steps_total;
    if (any_done) s->dirty = true;
  }
  hipError_t e = hipStreamSynchronize(hs);
  if (st == GD_OK && e != hipSuccess) st = hip_fail(e, "sweep");
  if (st != GD_OK) return st;
  // (4) convergence: a job's sweep ends every sweep_steps of its own steps
  if (s->cfg.converge_sweeps > 0) {
    for (ServJob *j : s->active) {
      if (j->converged || j->done_steps >= j->mc.steps_total || j->done_steps == 0 ||
          j->done_steps % s->cfg.sweep_steps != 0)
        continue;
      float b;
      if ((st = job_best(j->job, hs, &b)) != GD_OK) return st;
      j->best_hist.push_back(b);
      const int n = (int)j->best_hist.size(), w = s->cfg.converge_sweeps;
      if (n > w && j->best_hist[n - 1 - w] - b < s->cfg.converge_tol) j->converged = true;
    }
  }
  // (5) retire finished and converged jobs
  std::vector<ServJob *> done;
  for (auto it = s->active.begin(); it != s->active.end();) {
    ServJob *j = *it;
    if (j->done_steps >= j->mc.steps_total || j->converged) {
      done.push_back(j);
      it = s->active.erase(it);
      s->dirty = true;
    } else {
      ++it;
    }
  }
  for (ServJob *j : done) {
    const gd_status rs = read_out(j->job, j->out);
    const double rsteps = (double)j->job->h.R * j->done_steps;
    gd_job_destroy(j->job);
    std::lock_guard<std::mutex> lk(s->mu);
    auto h = s->held.find(j->ticket);
    if (h != s->held.end()) h->second->done = true;
    s->results.push_back({j->ticket, rs, j->t_submit, j->t_admit, s->now(), rsteps});
    s->used -= j->foot;
    delete j;
  }
  if (n_active) *n_active = (int)s->active.size();
  std::lock_guard<std::mutex> lk(s->mu);
  if (n_done) *n_done = (int)s->results.size();
  return GD_OK;
}
