/*
 * geauxdock_amd.h -- C ABI of the MI355X (gfx950) GeauxDock Monte Carlo hot path.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns a
 * gd_status; nothing calls exit().  The library owns device memory; callers own
 * every host array they pass in.
 *
 * Reference interfaces replaced (file:line in geaux-computational-bio/geauxdock):
 *   gd_job_create        <- CopyH2D / DeviceAlloc      src/backend_gpu/prepair.h:5,8
 *   gd_dock              <- Dock (GPU backend)          src/backend_gpu/backend_gpu.h:18-23
 *   gd_energy_eval       <- the s2max==1 initial-energy launch of MonteCarlo_d
 *                           src/backend_gpu/backend_gpu.cu:87-91 (kernel_cuda_l1_montecarlo.cu:29)
 *   gd_mc_param.seed/job_id <- InitCurand (per-replica cuRAND states are
 *                           replaced by a counter-based Philox stream keyed by
 *                           seed; no device state)   src/backend_gpu/backend_gpu.h:4
 *   gd_set_temperature   <- SetTemperature              src/frontend/util.C:325-363
 *   gd_load_complex      <- server1.C job assembly      src/frontend/server1.C:54-217
 *   gd_batch_dock        <- (no counterpart: many complexes in one launch sequence)
 *   gd_server_*          <- (no counterpart: continuous-batching request loop;
 *                           supersedes the MPI job farm src/main_mpi/server.C:27-66)
 * The reference's own C++ signatures (SetDevice, DeviceAlloc, ... Dock) are
 * exported too, see geauxdock_dropin.h.
 */
#ifndef GEAUXDOCK_AMD_H
#define GEAUXDOCK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Table dimensions of the force field (reference size.h:64-85). */
#define GD_NPT 31        /* protein effective-point types (MAXTP1) */
#define GD_NLT 24        /* ligand SYBYL atom types (MAXTP2) */
#define GD_NRES 20       /* residue types (MAXTP4) */
#define GD_NELE 50       /* electrostatic table size (MAXTP3) */
#define GD_NTERM 10      /* energy terms incl. total (MAXWEI) */
#define GD_MCS_INVALID 0.01234f  /* unmatched MCS position sentinel (size.h:107) */
#define GD_BADTYPE 50    /* unknown SYBYL type code (size.h:101) */

/* Energy term indices (reference geauxdock.h:131-147). */
enum { GD_VDW = 0, GD_ELE, GD_PMF, GD_PSP, GD_HDB, GD_HPC, GD_KDE, GD_LHM, GD_DST, GD_TOTAL };

typedef enum {
  GD_OK = 0,
  GD_ERR_ARG = 1,      /* NULL pointer, negative size, inconsistent arguments */
  GD_ERR_HIP = 2,      /* a HIP runtime call failed (gd_last_error() has the text) */
  GD_ERR_SIZE = 3,     /* size beyond a documented limit */
  GD_ERR_TYPE = 4,     /* unknown atom / point / residue type (e.g. BADKDE) */
  GD_ERR_IO = 5,       /* file missing or unparsable */
  GD_ERR_NOMEM = 6,    /* device or host allocation failed */
  GD_ERR_STATE = 7,    /* call out of order (e.g. no device initialised) */
  GD_ERR_BUSY = 8,     /* serving: admission refused (HBM budget) */
  GD_ERR_NODEV = 9     /* no usable gfx950 device */
} gd_status;

/* Derived force field, reference EnePara semantics (geauxdock_soa.h:154-187,
 * produced by OptimizeEnepara util_optimize_soa.C:489-540).  Tables are
 * indexed [ligand type][protein point type]. */
typedef struct {
  float p1a[GD_NLT][GD_NPT];
  float p2a[GD_NLT][GD_NPT];
  float lj0, lj1;
  float el1, el0, a1, b1;
  float pmf0[GD_NLT][GD_NPT];
  float pmf1[GD_NLT][GD_NPT];
  float hdb0[GD_NLT][GD_NPT];
  float hdb1[GD_NLT][GD_NPT];
  float hpp[GD_NRES];
  float hpl0[GD_NLT], hpl1[GD_NLT], hpl2[GD_NLT];
  float kde2, kde3;
  float a_para[GD_NTERM], b_para[GD_NTERM], w[GD_NTERM];
  float ele[GD_NELE];   /* raw electrostatic table, used by the protein loader */
} gd_enepara;

/* One docking complex ("job"): all arrays are host memory owned by the caller
 * and only read during gd_job_create.  Atom order is the reference's
 * type-sorted order (x3/t3/c3 of geauxdock_soa.h:30-34); conformers share the
 * atom types/charges, protein conformers share every point attribute. */
typedef struct {
  int n_lig, n_prt, n_tmp;                 /* ensemble sizes; n_rep = n_lig*n_prt*n_tmp */
  int lig_natom, prt_npoint, kde_npoint, mcs_nrow;
  /* ligand conformers [n_lig][lig_natom], relative to the conformer centroid */
  const float *lig_x, *lig_y, *lig_z;
  const int *lig_t;                        /* [lig_natom] SYBYL code 0..23 */
  const float *lig_c;                      /* [lig_natom] partial charge */
  const int *lig_perm;                     /* [lig_natom] file-order index of each atom, or NULL */
  float lig_center[3];                     /* native centroid in the lab frame */
  /* protein: coordinates [n_prt][prt_npoint]; attributes [prt_npoint] */
  const float *prt_x, *prt_y, *prt_z;
  const int *prt_t, *prt_c, *prt_cdc, *prt_seq3r;
  const float *prt_hpp, *prt_ele;
  float pocket_center[3];
  /* pocket-specific potential psp[ligand type][residue serial], [GD_NLT][psp_ncol] */
  const float *psp;
  int psp_ncol;
  /* KDE anchor points, sorted by type [kde_npoint] */
  const int *kde_t;
  const float *kde_x, *kde_y, *kde_z;
  /* MCS restraints [mcs_nrow][lig_natom] (same atom order as the ligand),
   * x == GD_MCS_INVALID marks an unmatched atom; tcc weight [mcs_nrow] */
  const float *mcs_x, *mcs_y, *mcs_z, *mcs_tcc;
  const gd_enepara *ene;
  const float *minus_beta;                 /* [n_tmp] = -1/T */
} gd_job_desc;

enum { GD_MOVE_REF_UNIFORM = 0,   /* reference: m += scale * U[0,1) (kernel_cpp_montecarlo_soa.C:248-249) */
       GD_MOVE_GAUSSIAN = 1 };    /* m += scale * N(0,1) */

/* Replica-exchange axes (gd_mc_param.exchange_axes bits; 0 means GD_EXCH_TEMP).
 * At every exchange point, in this order:
 *  - TEMP: neighbour temperatures swap labels inside each (ligand conformer,
 *    protein conformer) ladder;
 *  - LIG:  replicas at the same temperature and protein conformer on neighbour
 *    ligand conformers swap poses (Hamiltonian exchange: one extra energy
 *    evaluation of every replica at its partner's pose);
 *  - PRT:  the same along the protein-conformer axis.
 * The reference declares exchange (geauxdock.h:340-369) but implements none. */
enum { GD_EXCH_TEMP = 1, GD_EXCH_LIG = 2, GD_EXCH_PRT = 4 };

typedef struct {
  int steps_total;          /* MC steps after the initial-energy pass */
  int steps_per_dump;       /* steps per record segment (reference STEPS_PER_DUMP) */
  int steps_per_exchange;   /* exchange period in MC steps; 0 = off */
  int move_mode;            /* GD_MOVE_* */
  float move_scale[6];      /* tx ty tz rx ry rz */
  uint64_t seed;            /* Philox4x32-10 key */
  uint32_t job_id;          /* Philox counter word 2: the job's stream */
  int record;               /* 1 = keep per-dump accept records (drop-in Dock) */
  int exchange_axes;        /* GD_EXCH_* bits; 0 = GD_EXCH_TEMP */
} gd_mc_param;

/* Per-replica results (any pointer may be NULL).  Replica order is the
 * reference SetReplica order idx = n_tmp*n_lig*p + n_lig*t + l (util.C:380). */
typedef struct {
  float *movematrix;        /* [n_rep][6] last accepted pose */
  float *energy;            /* [n_rep][10] last accepted terms (normalised) + total */
  int *idx_tmp;             /* [n_rep] temperature index after exchanges */
  int *n_accept;            /* [n_rep] accepted MC moves */
  int *n_exchange;          /* [n_rep] accepted temperature swaps */
  float *best_energy;       /* [n_rep] lowest accepted total energy */
  int *is_accept;           /* [n_rep] last step (or conformer swap) accepted */
  int *n_swap_lig;          /* [n_rep] accepted pose swaps along the ligand-conformer axis */
  int *n_swap_prt;          /* [n_rep] accepted pose swaps along the protein-conformer axis */
} gd_replica_out;

typedef struct {
  double ms_total;          /* wall time of the MC loop (host clock) */
  double ms_pair_kernel;    /* summed HIP-event time of the timed pair-energy launches (on n
                               streams: of the n part launches of each timed step) */
  double ms_finish_kernel;  /* reserved, 0: the tail kernels are not event-timed */
  int n_pair_launch;        /* pair-kernel launches timed (steps 0, 16, 32, ...; n per step on
                               n streams) */
  double pair_evals;        /* ligand-atom x protein-point evaluations done */
  double replica_steps;     /* replica x MC steps done (initial pass excluded) */
  double ms_step_period;    /* steady-state MC step period: from the first to the last timed
                               step's pair launch start (HIP events on the pair kernel's
                               stream) over the steps between */
  double ms_pair_span;      /* mean time from the first start to the last end of a timed step's
                               pair launches (the launch itself on one stream) */
} gd_stats;

typedef struct gd_job gd_job;         /* device-resident complex */
typedef struct gd_server gd_server;   /* serving loop */

/* ---- device ---------------------------------------------------------- */
gd_status gd_init(int device);                 /* select + check a gfx950 device */
gd_status gd_device_info(char *name, int name_len, int *n_cu, int *clock_khz,
                         size_t *hbm_bytes);
const char *gd_last_error(void);
const char *gd_version(void);

/* ---- host helpers (reference frontend semantics) ---------------------- */
/* Geometric ladder T_i = floor * r^i (util.C:325-363); n == 1 uses floor. */
gd_status gd_set_temperature(int n_tmp, float floor_temp, float ceiling_temp,
                             float *minus_beta_out);

/* ---- jobs ------------------------------------------------------------- */
gd_status gd_job_create(const gd_job_desc *desc, gd_job **out);
gd_status gd_job_destroy(gd_job *job);
gd_status gd_job_sizes(const gd_job *job, int *n_rep, int *lig_natom, int *prt_npoint,
                       int *kde_npoint, int *mcs_nrow);
size_t gd_job_device_bytes(const gd_job *job);

/* Energy of arbitrary poses.  conf[n][2] = (ligand conformer, protein
 * conformer); pose[n][6] = movematrix.  terms[n][10] are the normalised
 * terms + total (what ReplicaMC.energy holds); raw[n][9] (optional) the
 * pre-normalisation sums vdw, ele, pmf, psp, -hdb/sqrt(2pi), hpc, kde/kde3,
 * lhm, dst; coords[n][lig_natom][3] (optional) the transformed ligand. */
gd_status gd_energy_eval(gd_job *job, int n_pose, const int *conf, const float *pose,
                         float *terms, float *raw, float *coords);

/* gd_energy_eval that allocates nothing: no hipMalloc, hipFree or other
 * allocating call, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  No counterpart in
 * the reference.  The kernels' scratch (partial sums, staged poses, work
 * items, results) lives in one static device array of
 * GD_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; poses run in
 * chunks of 64-pose blocks that fit it, and a job whose single block needs
 * more returns GD_ERR_SIZE.  Same results, bit for bit, as gd_energy_eval
 * (terms[n][10], raw[n][9] optional; no coordinates).  The job's replica
 * state is not touched.  Calls that overlap on one GPU share the array and
 * race; the array keeps the last call's contents. */
#define GD_NOALLOC_SCRATCH_BYTES (64 << 20)
gd_status gd_energy_eval_noalloc(gd_job *job, int n_pose, const int *conf, const float *pose,
                                 float *terms, float *raw);

/* Plain one-job docking: initial-energy pass, then steps_total MC steps with
 * temperature exchange, replica state resident on the device. */
gd_status gd_dock(gd_job *job, const gd_mc_param *mc, gd_replica_out *out, gd_stats *stats);

/* Current replica state of the job, read back from the device (the state
 * stays resident; the reference left the host copy stale after Dock,
 * backend_gpu.cu:64-133).  Same content as gd_dock's `out`. */
gd_status gd_job_replicas(gd_job *job, gd_replica_out *out);

/* Accept records of the last dump segment: rec[n_rep][cap] entries of
 * (step, movematrix[6], energy[10]) packed as 17 floats; count[n_rep].
 * A record holds a replica's accepted state at the start of step `step`,
 * kept when the step before it (or the initial pass, or a conformer swap)
 * was accepted (kernel_cpp_montecarlo_soa.C:224-231); at most
 * steps_per_dump per replica per dump (the reference's Record,
 * geauxdock.h:374-378).  GD_ERR_STATE when no dump has recorded. */
gd_status gd_dock_records(gd_job *job, int cap, float *rec, int *count);

/* Receiver of one dump's records (gd_dock_stream): dump `dump` covers MC
 * steps [step_begin, step_begin + cap); rec points at n_rep rows of
 * cap * 17 floats (the gd_dock_records layout with cap = steps_per_dump),
 * count[n_rep] gives the entries used in each row.  Called on the thread that
 * runs gd_dock_stream, once per dump and in dump order; the arrays are valid
 * only during the call, and the callback must not call the library on the
 * same job. */
typedef void (*gd_record_fn)(void *user, int dump, int step_begin, int n_rep, int cap,
                             const float *rec, const int *count);
enum { GD_STREAM_BLOCKING = 1 };
/* gd_dock with every dump's accept records handed to fn (records on
 * whatever mc->record says).  By default each dump's records are copied to
 * pinned host memory on a second stream while the device runs the next dump;
 * GD_STREAM_BLOCKING waits for the dump and copies synchronously instead, the
 * reference's per-dump cudaMemcpy (backend_gpu.cu:96-118) -- the plain path in
 * the reference's shape. */
gd_status gd_dock_stream(gd_job *job, const gd_mc_param *mc, gd_record_fn fn, void *user, int flags,
                         gd_replica_out *out, gd_stats *stats);

/* A ready-made record sink: appends PrintRecord's lines (util_print.C:70-151:
 * "%3d %d" replica and step, then " %.4f" per energy term) for replicas
 * rep_lo..rep_hi of every dump to a text file; header != 0 writes the
 * "rep step vdw ... total" title line first.  Pass gd_record_csv_write as
 * the gd_record_fn and the writer as its user pointer. */
typedef struct gd_record_csv gd_record_csv;
gd_status gd_record_csv_open(const char *path, int rep_lo, int rep_hi, int header,
                             gd_record_csv **out);
void gd_record_csv_write(void *writer, int dump, int step_begin, int n_rep, int cap,
                         const float *rec, const int *count);
/* Closes the file; returns GD_ERR_IO if any write failed. */
gd_status gd_record_csv_close(gd_record_csv *writer);

/* Many complexes in one launch sequence (persistent work-queue pair kernel,
 * graph-captured sweep segments).  mc[i] / out[i] per job. */
gd_status gd_batch_dock(gd_job **jobs, int n_jobs, const gd_mc_param *mc,
                        gd_replica_out *out, gd_stats *stats);

/* ---- tiling / autotuner ------------------------------------------------ */
typedef struct {
  int atoms_per_lane;       /* ligand atoms per work item (4, 8 or 16) */
  int tile_points;          /* protein points per tile */
  int waves_per_cu;         /* persistent grid size per CU */
} gd_tiling;
gd_status gd_tiling_get(const gd_job *job, gd_tiling *t);
/* Streams gd_dock steps the job on (1 to 4, default 2): with n, each steps one
 * of n contiguous parts of the replica blocks between exchange points, so one
 * part's reduce and finish kernels and launch tail overlap the other parts'
 * pair work.  The results are the same bits either way.  Waits for the job's
 * stream (the parts' work lists are rebuilt). */
gd_status gd_job_set_streams(gd_job *job, int n_streams);
gd_status gd_tiling_set(gd_job *job, const gd_tiling *t);
/* Sweep atoms per lane x protein tile counts on `job` for `steps` MC steps
 * each (33 or more for an event-timed period) on the job's streams, and apply
 * the one with the shortest MC step period (*best_ms) to this job only (the
 * reference's counterpart, the disabled CalcThreadBlockShape of
 * backend_gpu/kernel_cuda_l2_util.cu:40-67, is a formula). */
gd_status gd_autotune(gd_job *job, int steps, gd_tiling *best, double *best_ms);
/* Sweep the tiling of shared launches (protein tile points x persistent-grid
 * waves per CU, 4 atoms per lane) on a batch of these jobs, `steps` MC steps
 * per candidate; returns the fastest.  Nothing process-wide changes: the
 * library's tilings come from its tuned table (geauxdock_amd/tune.py writes
 * it from these sweeps). */
gd_status gd_autotune_shared(gd_job **jobs, int n_jobs, int steps, gd_tiling *best, double *best_ms);
/* The tiling the library gives a job of these sizes, one job per launch
 * (shared == 0) or in shared launches (batches, serving). */
gd_status gd_tiling_lookup(int lig_natom, int prt_npoint, int n_rep, int shared, gd_tiling *t);

/* ---- serving ------------------------------------------------------------ */
typedef struct {
  size_t hbm_budget;        /* bytes of device memory jobs may occupy */
  int max_active;           /* cap on concurrently active jobs */
  int sweep_steps;          /* MC steps per sweep between admissions */
  /* Convergence retirement (0 = off): a job retires before its step count
   * once the lowest energy any of its replicas has accepted improved by less
   * than converge_tol over its last converge_sweeps sweeps.  A job's sweeps
   * end every sweep_steps of its own steps, so the decision depends on the
   * job alone, not on what else is active. */
  int converge_sweeps;
  float converge_tol;
} gd_server_config;

typedef struct {
  int64_t ticket;
  int status;               /* gd_status of the job */
  double t_submit, t_admit, t_done;   /* seconds, server clock */
  double replica_steps;
} gd_job_result;

/* A job's result: status GD_OK; replica_steps counts the steps it ran (fewer
 * than requested when it retired on convergence).  gd_server_step and
 * gd_server_cancel of active jobs are serialised with each other; submits,
 * results and usage queries never wait for a device sweep. */
gd_status gd_server_create(const gd_server_config *cfg, gd_server **out);
gd_status gd_server_destroy(gd_server *s);
/* Device bytes a complex will occupy as a plain job (host-side estimate of
 * what gd_job_create allocates with the default tiling).  The server admits
 * by the same estimate for its shared-launch tiling and then charges the
 * job's actual gd_job_device_bytes. */
size_t gd_job_footprint(const gd_job_desc *desc);
/* Queue a docking request.  The descriptor's arrays must stay valid until the
 * request is admitted (its results are ready or it was cancelled); out
 * (optional) receives the final replica state when the job retires.  The
 * server creates the device job at admission, when the job's footprint fits
 * the HBM budget and an active slot is free. */
gd_status gd_server_submit(gd_server *s, const gd_job_desc *desc, const gd_mc_param *mc,
                           gd_replica_out *out, int64_t *ticket);
/* One sweep: admit what fits, advance every active job by up to sweep_steps
 * MC steps in shared launches, retire finished jobs.  *n_active / *n_done
 * receive the active count and the number of results ready to pop. */
gd_status gd_server_step(gd_server *s, int *n_active, int *n_done);
gd_status gd_server_cancel(gd_server *s, int64_t ticket);
/* Pop up to `cap` finished results. */
gd_status gd_server_results(gd_server *s, gd_job_result *res, int cap, int *n);
double gd_server_clock(gd_server *s);
/* Bytes of the budget in use and requests still waiting for admission
 * (file requests still being parsed included). */
gd_status gd_server_usage(gd_server *s, size_t *bytes_used, int *n_pending);

/* A docking request given by its files: the inputs of gd_load_complex. */
typedef struct {
  const char *sdf, *ff, *pdb, *ligand_id;   /* copied at submission */
  const gd_enepara *ene;                    /* must stay valid until the request is parsed */
  int n_tmp;
  float floor_temp, ceiling_temp;
} gd_request_files;
/* Queue a request by its files.  The server's loader thread parses them (the
 * host input pipeline of server1.C:54-217: ligand, restraint and protein
 * loaders with their optimisation passes) while the device runs the active
 * jobs; the parsed job then waits for admission like a gd_server_submit
 * request.  The server keeps its final replica state until gd_server_take.
 * A request that fails to parse, or can never fit the budget, comes back as a
 * result with that status. */
gd_status gd_server_submit_files(gd_server *s, const gd_request_files *req, const gd_mc_param *mc,
                                 int64_t *ticket);
/* Final replica state of a finished file request.  With out == NULL, *n_rep
 * receives its replica count and the state stays; with out (arrays of *n_rep
 * entries; any pointer may be NULL) it is copied out and released.
 * GD_ERR_STATE while the job has not finished, GD_ERR_ARG for an unknown
 * ticket. */
gd_status gd_server_take(gd_server *s, int64_t ticket, int *n_rep, gd_replica_out *out);

/* ---- loaders (reference frontend, re-implemented) ----------------------- */
typedef struct gd_host_job gd_host_job;   /* host-side complex built by the loaders */
/* Loads paras + 08ff_opt + 08_nor_a/b (loadEnePara/loadWeight/loadNorPara +
 * OptimizeEnepara, load.C:1189-1507, util_optimize_soa.C:489-540). */
gd_status gd_load_enepara(const char *paras, const char *weight, const char *nor_a,
                          const char *nor_b, gd_enepara *out);
/* One ligand .sdf + its eFindSite .ff + protein .pdb, as server1.C builds a
 * Complex: loadLigand/loadLHM/loadProtein + Optimize*.  n_tmp temperatures
 * on the given ladder. */
gd_status gd_load_complex(const char *sdf, const char *ff, const char *pdb,
                          const char *ligand_id, const gd_enepara *ene, int n_tmp,
                          float floor_temp, float ceiling_temp, gd_host_job **out);
/* Synthetic complex (seeded, deterministic): prt_npoint effective points,
 * lig_natom heavy atoms, ensemble sizes as given. */
gd_status gd_synth_complex(int prt_npoint, int lig_natom, int n_lig, int n_prt, int n_tmp,
                           float floor_temp, float ceiling_temp, uint64_t seed,
                           const gd_enepara *ene, gd_host_job **out);
/* One complex of a virtual screen: the protein target of protein_seed (the
 * same points and pocket for every ligand) with the ligand, KDE anchors, MCS
 * rows and psp entries of ligand_seed. */
gd_status gd_synth_screen(int prt_npoint, uint64_t protein_seed, int lig_natom, uint64_t ligand_seed,
                          int n_lig, int n_prt, int n_tmp, float floor_temp, float ceiling_temp,
                          const gd_enepara *ene, gd_host_job **out);
/* View of a host job as a gd_job_desc (pointers stay valid until free). */
gd_status gd_host_job_desc(const gd_host_job *h, gd_job_desc *desc);
gd_status gd_host_job_free(gd_host_job *h);
/* Conformer RMSD after optimal superposition (replaces rmsd.f u3b, load.C:421). */
float gd_kabsch_rmsd(int n, const float *a_xyz, const float *b_xyz);

/* ---- diagnostics (tests and measurement; no reference counterpart) ------ */
/* Philox4x32-R of one counter block (known-answer tests). */
void gd_philox(const uint32_t ctr[4], const uint32_t key[2], int rounds, uint32_t out[4]);
/* Move increments as gd_dock draws them: inc[s][r][6] for replicas r < n_rep
 * at MC steps step0 + s, s < n_steps (at most 2^28 replica-steps), computed by
 * the device kernels' code (on_device != 0) or by the same source compiled for
 * the host.  The reference move it generalises: kernel_cpp_montecarlo_soa.C:242-250. */
gd_status gd_move_draws(uint64_t seed, uint32_t job_id, int move_mode, const float move_scale[6],
                        int n_rep, uint32_t step0, int n_steps, int on_device, float *inc);
/* Increments from given Philox words: w[n][8] -> inc[n][6] (host code path). */
gd_status gd_move_incr_words(int move_mode, const float move_scale[6], int n, const uint32_t *w,
                             float *inc);
/* Engine clock under an all-CU FMA load, in MHz, measured over about `ms`
 * milliseconds (1..2000) from the shader-clock and constant-rate counters. */
gd_status gd_measure_sclk(int ms, double *sclk_mhz);
/* A plain device-to-device copy of `bytes` (a multiple of 16), `reps` times,
 * by a grid-stride 16-byte vector kernel: the known-size calibration of HBM
 * traffic counters.  *ms (optional): mean milliseconds per copy. */
gd_status gd_device_copy(size_t bytes, int reps, double *ms);
/* The pair kernel's far-field and near-branch r^2 thresholds of a
 * (ligand type, point type) parameter row (tests of the exact-zero shortcut). */
float gd_far_r2(float pmf0, float el1);
float gd_near_r2(float pmf0, float el1, float hdb0, float hdb1);

#ifdef __cplusplus
}
#endif
#endif
