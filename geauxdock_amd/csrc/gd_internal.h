// gd_internal.h -- host-side job object and runtime helpers shared by the
// runtime translation units (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/geauxdock_amd.h"
#include "gd_device.h"
#include "gd_rng.h"

struct gd_job {
  GdJob h{};                 // host mirror of the device struct
  GdJob *d_job = nullptr;    // device copy
  void *d_base = nullptr;    // tables + replica state
  void *d_part = nullptr;    // partial sums (tiling-dependent)
  size_t part_bytes = 0;
  void *d_lists = nullptr;   // work items + blocks
  size_t lists_bytes = 0;
  void *d_rec = nullptr;     // accept records: a two-half ring, one half per dump
  size_t rec_bytes = 0;
  void *h_rec = nullptr;     // pinned host copies of both halves (record streaming)
  size_t h_rec_bytes = 0;
  hipStream_t copy_stream = nullptr;  // D2H of finished dumps beside the MC stream
  int last_half = -1;        // ring half of the last completed dump (-1: none)
  size_t bytes = 0;          // d_base size
  std::vector<GdItem> items; // job-local work items (job index 0)
  GdItem *d_items = nullptr;
  GdBlock *d_blocks = nullptr;
  int n_items = 0, n_blocks = 0;
  // the plain path's stepping parts: n_parts contiguous ranges of the blocks,
  // [b0_part[k], b0_part[k] + nb_part[k]), each with its items (the job's items
  // of those blocks, in the job's order) and a stream (part 0: the job's own)
  static constexpr int kMaxParts = 4;
  GdItem *d_items_part[kMaxParts] = {};
  int n_items_part[kMaxParts] = {}, nb_part[kMaxParts] = {}, b0_part[kMaxParts] = {};
  int n_parts = 0;
  int n_streams = 2;         // streams the plain path steps on (1 .. kMaxParts)
  hipStream_t part_stream[kMaxParts] = {};
  // protein points in source order, laid out on the device per tile count
  std::vector<float4> src_pco;     // [n_prt][P]
  std::vector<float4> src_pattr;   // [P]
  std::vector<float> src_pspv;     // [n_types + 1][P]
  std::vector<int> src_key;        // [P] class * GD_PTAB_PT + point type
  std::vector<int> src_spread;     // [P] order inside a (class, type) group
  std::vector<int> atom_kde;       // [L] KDE anchors of each device atom (item order)
  int laid_nP = 0;                 // tile count of the device order (0: none yet)
  hipStream_t stream = nullptr;
};

namespace gdr {
void set_error(const std::string &s);
gd_status hip_fail(hipError_t e, const char *what);
gd_status ensure_device();
int device();   // the device the library runs on (-1 before initialisation)
int n_cu();
void default_tiling(int L, int P, int R, int &lc, int &ptile);  // one job per launch
int shared_ptile(int L, int P);      // tile points in shared launches (4 atoms per lane)
int pair_waves_per_cu(int lc);       // persistent pair kernel residency
int lg2_floor(int x);
gd_status apply_tiling(gd_job *j, int lc, int ptile);
gd_status sync_job_struct(gd_job *j);
gd_status reset_replicas(gd_job *j);
void set_mc(GdJob &g, const gd_mc_param *mc);
gd_status read_out(gd_job *j, gd_replica_out *out);
float far_r2_threshold(double pmf0, double el1);
gd_status launch_pair(int lc, const GdJob *d_jobs, const GdItem *d_items, int n_items, hipStream_t s);
gd_status launch_finish(const GdJob *d_jobs, const GdBlock *d_blocks, int n_blocks, int mode,
                        int step, hipStream_t s);
gd_status launch_transform(const GdJob *d_jobs, const GdBlock *d_blocks, int n_blocks, hipStream_t s);
gd_status launch_exchange(const GdJob *d_jobs, const GdExch *d_ex, int n_ex, hipStream_t s);
// One conformer swap sub-step along `axis` (0 ligand, 1 protein) for the jobs
// of the lists whose swr[job] round is >= 0: propose, pair (the caller's
// launcher), reduce, finish(EVAL), accept.  launch_step / step_base as for the
// finish kernel of the step just completed; wq is reset for the next pair
// launch when given.
gd_status launch_swap(int axis, const GdJob *d_jobs, const GdBlock *d_blocks, int n_blocks,
                      const int *d_swr, int launch_step, const int *step_base, int *wq,
                      const std::function<gd_status()> &pair, hipStream_t s);

// HIP-event timer for one kernel over many launches: events are recorded
// around the timed launches and summed at synchronisation points, so timing
// never stalls the launch queue.  Each record still costs the queue about
// 4 us between the kernels around it, so an MC loop times the pair launch of
// one step in kTimedEvery (step 0, kTimedEvery, ...): the mean launch time
// comes from that sample at well under 1% of the loop's time.
constexpr int kTimedEvery = 16;
struct KTimer {
  std::vector<hipEvent_t> ev;
  int used = 0;
  double ms = 0.0;
  int n = 0;
  double first_to_last_ms = 0.0;  // first to last timed launch start (collect)
  ~KTimer();
  gd_status begin(hipStream_t s);   // records start event
  gd_status end(hipStream_t s);     // records stop event
  gd_status collect();              // after a stream sync
};
}  // namespace gdr
