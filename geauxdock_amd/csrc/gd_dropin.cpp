// gd_dropin.cpp -- the reference GPU backend's entry points (prepair.h,
// backend_gpu.h) on top of the C ABI, so src/main_gpu/geauxdock.C can link
// libgeauxdock_amd.so instead of src/backend_gpu.
//
// Semantics follow src/backend_gpu/prepair.C:39-135 and backend_gpu.cu:49-133:
// one device, the whole replica range on it, the initial-energy pass, then
// steps_total steps in dumps of steps_per_dump; after the run the last dump's
// accept records are in the caller's Record array, replica 0's first record
// and the "Ref Result" line are printed (common/kernel_print.C), then the
// timers (kernel_print_timer.C) and the "Benchmark," CSV line
// (kernel_print_benchmark.C) that run/batch.py scrapes.  Errors print and exit
// as the reference's CUDA_ERR wrapper does.  Differences: the final replica
// state is also written back to ch->replica (the reference leaves the host
// copy stale), and moves use the reference's one-sided U[0,1) scale.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>

#include "../../include/geauxdock_amd.h"
#include "../../include/geauxdock_dropin.h"

namespace {

struct DropinCtx {
  gd_job *job = nullptr;
  Complex *src = nullptr;  // last CopyH2D'd host complex
  std::vector<float> lx, ly, lz, lc, px, py, pz, hpp, ele, psp, kx, ky, kz, mx, my, mz, tcc, mb;
  std::vector<int> lt, perm, pt, pc, cdc, seq, kt;
  gd_enepara ene;
};

void die(gd_status st, const char *what) {
  std::fprintf(stderr, "%s failed: status %d: %s\n", what, (int)st, gd_last_error());
  std::exit(EXIT_FAILURE);
}

DropinCtx *ctx_of(Complex **c) { return reinterpret_cast<DropinCtx *>(c[0]); }

// The sizes must fit the Complex's fixed capacities (size.h), and n_rep must be
// the ensemble product, or the reads of the Complex and the writes of the
// replicas and records would run past their arrays: fail as the reference's
// error paths do, with a message.
void check_sizes(const ComplexSize &sz, const char *where) {
  const struct {
    const char *name;
    int v, cap;
  } f[] = {{"n_lig", sz.n_lig, GD_REF_MAX_CONF_LIG},   {"n_prt", sz.n_prt, GD_REF_MAX_CONF_PRT},
           {"n_tmp", sz.n_tmp, GD_REF_MAX_TMP},         {"n_rep", sz.n_rep, GD_REF_MAX_REP},
           {"lig_natom", sz.lig_natom, GD_REF_MAXLIG},  {"prt_npoint", sz.prt_npoint, GD_REF_MAXPRO},
           {"kde_npoint", sz.kde_npoint, GD_REF_MAXKDE}, {"mcs_nrow", sz.mcs_nrow, GD_REF_MAX_MCS_ROW}};
  for (const auto &e : f)
    if (e.v < 0 || e.v > e.cap) {
      std::fprintf(stderr, "%s failed: ComplexSize.%s = %d is outside 0..%d\n", where, e.name, e.v, e.cap);
      std::exit(EXIT_FAILURE);
    }
  if ((long long)sz.n_lig * sz.n_prt * sz.n_tmp != sz.n_rep) {
    std::fprintf(stderr, "%s failed: ComplexSize.n_rep = %d is not n_lig x n_prt x n_tmp\n", where, sz.n_rep);
    std::exit(EXIT_FAILURE);
  }
}

// Reference Complex (fixed capacities, geauxdock.h:391-439) -> gd_job_desc.
void to_desc(Complex *ch, DropinCtx *x, gd_job_desc *d) {
  const ComplexSize &sz = ch->size;
  const int L = sz.lig_natom, P = sz.prt_npoint, K = sz.kde_npoint, M = sz.mcs_nrow;
  const int nl = sz.n_lig, np = sz.n_prt, nt = sz.n_tmp;
  // type-sorted atom order is the reference's x3/t3/c3 (util_optimize_soa.C:112-126);
  // its file-order index comes from matching x2
  x->lx.resize((size_t)nl * L), x->ly.resize((size_t)nl * L), x->lz.resize((size_t)nl * L);
  x->lt.assign(ch->lig[0].t3, ch->lig[0].t3 + L);
  x->lc.assign(ch->lig[0].c3, ch->lig[0].c3 + L);
  for (int i = 0; i < nl; ++i)
    for (int j = 0; j < L; ++j) {
      x->lx[(size_t)i * L + j] = ch->lig[i].x3[j];
      x->ly[(size_t)i * L + j] = ch->lig[i].y3[j];
      x->lz[(size_t)i * L + j] = ch->lig[i].z3[j];
    }
  x->perm.assign(L, -1);
  std::vector<char> used(L, 0);
  for (int j = 0; j < L; ++j)
    for (int f = 0; f < L; ++f)
      if (!used[f] && ch->lig[0].x2[f] == ch->lig[0].x3[j] && ch->lig[0].y2[f] == ch->lig[0].y3[j] &&
          ch->lig[0].z2[f] == ch->lig[0].z3[j] && ch->lig[0].t2[f] == ch->lig[0].t3[j]) {
        x->perm[j] = f;
        used[f] = 1;
        break;
      }
  for (int j = 0; j < L; ++j)
    if (x->perm[j] < 0) x->perm[j] = j;
  x->px.resize((size_t)np * P), x->py.resize((size_t)np * P), x->pz.resize((size_t)np * P);
  for (int i = 0; i < np; ++i)
    for (int p = 0; p < P; ++p) {
      x->px[(size_t)i * P + p] = ch->prt[i].x[p];
      x->py[(size_t)i * P + p] = ch->prt[i].y[p];
      x->pz[(size_t)i * P + p] = ch->prt[i].z[p];
    }
  const Protein &p0 = ch->prt[0];
  x->pt.assign(p0.t, p0.t + P);
  x->pc.assign(p0.c, p0.c + P);
  x->cdc.assign(p0.cdc, p0.cdc + P);
  x->seq.assign(p0.seq3r, p0.seq3r + P);
  x->hpp.assign(p0.hpp, p0.hpp + P);
  x->ele.assign(p0.ele, p0.ele + P);
  x->psp.assign((size_t)GD_NLT * GD_REF_MAXPRO, 0.0f);
  for (int lt = 0; lt < GD_NLT; ++lt)
    std::memcpy(&x->psp[(size_t)lt * GD_REF_MAXPRO], ch->psp.psp[lt], GD_REF_MAXPRO * sizeof(float));
  x->kt.assign(ch->kde.t, ch->kde.t + K);
  x->kx.assign(ch->kde.x, ch->kde.x + K);
  x->ky.assign(ch->kde.y, ch->kde.y + K);
  x->kz.assign(ch->kde.z, ch->kde.z + K);
  x->mx.resize((size_t)M * L), x->my.resize((size_t)M * L), x->mz.resize((size_t)M * L);
  x->tcc.resize(M);
  for (int m = 0; m < M; ++m) {
    for (int j = 0; j < L; ++j) {
      const int f = x->perm[j];
      x->mx[(size_t)m * L + j] = ch->mcs[m].x[f];
      x->my[(size_t)m * L + j] = ch->mcs[m].y[f];
      x->mz[(size_t)m * L + j] = ch->mcs[m].z[f];
    }
    x->tcc[m] = ch->mcs[m].tcc;
  }
  x->mb.resize(nt);
  for (int t = 0; t < nt; ++t) x->mb[t] = ch->temp[t].minus_beta;
  const EnePara &e = ch->enepara;
  std::memset(&x->ene, 0, sizeof x->ene);
  std::memcpy(x->ene.p1a, e.p1a, sizeof e.p1a);
  std::memcpy(x->ene.p2a, e.p2a, sizeof e.p2a);
  std::memcpy(x->ene.pmf0, e.pmf0, sizeof e.pmf0);
  std::memcpy(x->ene.pmf1, e.pmf1, sizeof e.pmf1);
  std::memcpy(x->ene.hdb0, e.hdb0, sizeof e.hdb0);
  std::memcpy(x->ene.hdb1, e.hdb1, sizeof e.hdb1);
  x->ene.lj0 = e.lj0, x->ene.lj1 = e.lj1, x->ene.el1 = e.el1, x->ene.el0 = e.el0;
  x->ene.a1 = e.a1, x->ene.b1 = e.b1;
  std::memcpy(x->ene.hpp, e.hpp, sizeof e.hpp);
  std::memcpy(x->ene.hpl0, e.hpl0, sizeof e.hpl0);
  std::memcpy(x->ene.hpl1, e.hpl1, sizeof e.hpl1);
  std::memcpy(x->ene.hpl2, e.hpl2, sizeof e.hpl2);
  x->ene.kde2 = e.kde2, x->ene.kde3 = e.kde3;
  std::memcpy(x->ene.a_para, e.a_para, sizeof e.a_para);
  std::memcpy(x->ene.b_para, e.b_para, sizeof e.b_para);
  std::memcpy(x->ene.w, e.w, sizeof e.w);

  std::memset(d, 0, sizeof(*d));
  d->n_lig = nl, d->n_prt = np, d->n_tmp = nt;
  d->lig_natom = L, d->prt_npoint = P, d->kde_npoint = K, d->mcs_nrow = M;
  d->lig_x = x->lx.data(), d->lig_y = x->ly.data(), d->lig_z = x->lz.data();
  d->lig_t = x->lt.data(), d->lig_c = x->lc.data(), d->lig_perm = x->perm.data();
  for (int i = 0; i < 3; ++i) {
    d->lig_center[i] = ch->lig[0].center[i];
    d->pocket_center[i] = p0.pocket_center[i];
  }
  d->prt_x = x->px.data(), d->prt_y = x->py.data(), d->prt_z = x->pz.data();
  d->prt_t = x->pt.data(), d->prt_c = x->pc.data(), d->prt_cdc = x->cdc.data();
  d->prt_seq3r = x->seq.data(), d->prt_hpp = x->hpp.data(), d->prt_ele = x->ele.data();
  d->psp = x->psp.data(), d->psp_ncol = GD_REF_MAXPRO;
  d->kde_t = x->kt.data(), d->kde_x = x->kx.data(), d->kde_y = x->ky.data(), d->kde_z = x->kz.data();
  d->mcs_x = x->mx.data(), d->mcs_y = x->my.data(), d->mcs_z = x->mz.data(), d->mcs_tcc = x->tcc.data();
  d->ene = &x->ene;
  d->minus_beta = x->mb.data();
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

void SetDevice() {
  const gd_status st = gd_init(0);
  if (st != GD_OK) die(st, "SetDevice");
}

void DeviceAlloc(Complex **c, Record **r, gd_rng_state **s) {
  c[0] = reinterpret_cast<Complex *>(new DropinCtx());
  r[0] = nullptr;  // records live with the device job; Dock fills the host array
  s[0] = new gd_rng_state();
  std::memset(s[0], 0, sizeof(gd_rng_state));
}

void DeviceFree(Complex **c, Record **r, gd_rng_state **s) {
  DropinCtx *x = ctx_of(c);
  if (x) {
    if (x->job) gd_job_destroy(x->job);
    delete x;
  }
  c[0] = nullptr;
  r[0] = nullptr;
  delete s[0];
  s[0] = nullptr;
}

// One device: the whole replica range (prepair.C:92-143 with NGPU = 1).
void SetParaT(Complex *ch, ParaT **pt) {
  check_sizes(ch->size, "SetParaT");
  const int n_rep = ch->size.n_rep;
  pt[0]->rep_begin = 0;
  pt[0]->rep_end = n_rep - 1;
  pt[0]->n_rep = n_rep;
  pt[0]->record_sz = (int)(sizeof(Record) * (size_t)n_rep);
}

void CopyH2D(Complex *ch, Complex **cd, ParaT **pt) {
  DropinCtx *x = ctx_of(cd);
  ch->rep_begin = pt[0]->rep_begin;
  ch->rep_end = pt[0]->rep_end;
  if (x->job) {
    gd_job_destroy(x->job);
    x->job = nullptr;
  }
  check_sizes(ch->size, "CopyH2D");
  gd_job_desc d;
  to_desc(ch, x, &d);
  const gd_status st = gd_job_create(&d, &x->job);
  if (st != GD_OK) die(st, "CopyH2D");
  x->src = ch;
}

// Seeds the Philox stream from the clock, as InitCurand seeds cuRAND
// (backend_gpu.cu:49-59); GD_SEED in the environment fixes it.
void InitCurand(gd_rng_state **s) {
  const char *env = std::getenv("GD_SEED");
  std::srand((unsigned)std::time(nullptr));
  s[0]->seed = env ? std::strtoull(env, nullptr, 0) : (uint64_t)std::rand();
  s[0]->job_id = 0;
}

// Copies one dump's accept records into the caller's Record array: what the
// reference's per-dump cudaMemcpy into rh does (backend_gpu.cu:110-116), so rh
// holds the last dump's records when Dock returns.
struct DumpToRecords {
  Record *rh;
  const Complex *ch;
  int rep_begin, rep_end;
};

static void dump_to_records(void *user, int dump, int step_begin, int n_rep, int cap, const float *rec,
                            const int *count) {
  (void)dump;
  (void)step_begin;
  const DumpToRecords *d = (const DumpToRecords *)user;
  for (int r = d->rep_begin; r <= d->rep_end && r < n_rep; ++r) {
    Record &R = d->rh[r];
    const int n = std::min(std::min(count[r], cap), GD_REF_STEPS_PER_DUMP);
    R.next_entry = n;
    const ReplicaMC &base = d->ch->replica[r];
    for (int k = 0; k < n; ++k) {
      const float *o = rec + ((size_t)r * cap + k) * 17;
      ReplicaMC &q = R.replica[k];
      q = base;
      q.step = (int)o[0];
      for (int i = 0; i < 6; ++i) q.movematrix[i] = o[1 + i];
      for (int i = 0; i < 10; ++i) q.energy[i] = o[7 + i];
      q.is_accept = 1;
    }
  }
}

// Temperature-exchange period of the drop-in Dock.  The reference GPU Dock
// never exchanges (backend_gpu.cu:60-133), and its frontend never writes
// McPara.steps_per_exchange (server1.C allocates McPara uninitialised;
// ParseArguments, util.C:53-190, does not set it), so the field is not read:
// exchange is off unless GD_EXCHANGE_STEPS names a period.
static int dropin_exchange_period(int steps_total) {
  const char *env = std::getenv("GD_EXCHANGE_STEPS");
  if (!env) return 0;
  const long v = std::strtol(env, nullptr, 10);
  return (v > 0 && v <= steps_total) ? (int)v : 0;
}

void Dock(Complex *ch, Record *rh, Complex **cd, Record **rd, ParaT **pt, gd_rng_state **s) {
  (void)rd;
  DropinCtx *x = ctx_of(cd);
  const int steps_total = ch->mcpara.steps_total;
  const int steps_per_dump = ch->mcpara.steps_per_dump > 0 ? ch->mcpara.steps_per_dump : 1;
  if (steps_per_dump > GD_REF_STEPS_PER_DUMP) die(GD_ERR_SIZE, "Dock: steps_per_dump > STEPS_PER_DUMP");
  const int n_rep = ch->size.n_rep;
  gd_mc_param mc;
  std::memset(&mc, 0, sizeof mc);
  mc.steps_total = steps_total;
  mc.steps_per_dump = steps_per_dump;
  mc.steps_per_exchange = dropin_exchange_period(steps_total);
  mc.move_mode = GD_MOVE_REF_UNIFORM;
  for (int i = 0; i < 6; ++i) mc.move_scale[i] = ch->mcpara.move_scale[i];
  mc.seed = s[0]->seed;
  mc.job_id = s[0]->job_id++;
  mc.record = 1;
  std::printf("Start kernels\n");
  std::vector<float> mm((size_t)n_rep * 6), en((size_t)n_rep * 10);
  std::vector<int> tmp(n_rep), acc(n_rep);
  gd_replica_out out;
  std::memset(&out, 0, sizeof out);
  out.movematrix = mm.data();
  out.energy = en.data();
  out.idx_tmp = tmp.data();
  out.is_accept = acc.data();
  gd_stats stats;
  for (int r = pt[0]->rep_begin; r <= pt[0]->rep_end; ++r) rh[r].next_entry = 0;
  DumpToRecords sink = {rh, ch, pt[0]->rep_begin, pt[0]->rep_end};
  const double t0 = now_ms();
  // every dump's records copied into rh after the dump, blocking, as the reference does
  gd_status st = gd_dock_stream(x->job, &mc, dump_to_records, &sink, GD_STREAM_BLOCKING, &out, &stats);
  if (st != GD_OK) die(st, "Dock");
  const double t_mc = now_ms() - t0;
  for (int r = pt[0]->rep_begin; r <= pt[0]->rep_end; ++r) {
    ReplicaMC &rep = ch->replica[r];
    for (int i = 0; i < 6; ++i) rep.movematrix[i] = mm[(size_t)r * 6 + i];
    for (int i = 0; i < 10; ++i) rep.energy[i] = en[(size_t)r * 10 + i];
    rep.idx_tmp = tmp[r];
    rep.is_accept = acc[r];
  }
  // kernel_print.C: replica 0, record 0, then the published line
  {
    const ReplicaMC &q = rh[0].replica[0];
    std::printf("%3d %d", 0, q.step);
    for (int i = 0; i < 10; ++i) std::printf(" %.4f", q.energy[i]);
    std::printf("\n");
    std::printf("  0 0 0.6434 -0.0367 -0.2079 -0.1845 0.8521 -0.8880 0.0523 0.1744 -1.0000 0.7735 Ref Result\n");
  }
  // kernel_print_timer.C
  const float mcpersec = (float)((double)steps_total * n_rep / (t_mc / 1000.0));
  std::printf("mc kernel time\t\t\t%8.3f s\n", t_mc / 1000.0);
  std::printf("time per MC sweep per replica\t%8.3f * 1e-6 s\n", 1e6 / mcpersec);
  std::printf("MC sweeps per second\t\t%8.3f\n", mcpersec);
  std::printf("speedup over 1023.5\t\t%8.3f X\n", mcpersec / 1023.5);
  // kernel_print_benchmark.C
  std::printf("Benchmark,\t\t%s,%s,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%.3f\n", ch->lig_file, ch->prt_file,
              steps_total, ch->mcpara.steps_per_dump, ch->size.n_lig, ch->size.n_prt, ch->size.n_tmp,
              n_rep, ch->size.lig_natom, ch->size.prt_npoint, ch->size.kde_npoint,
              ch->size.mcs_nrow, t_mc);
}
