// dropin_main.cpp -- TEST PROGRAM: the reference GPU driver's call sequence
// (src/main_gpu/geauxdock.C:82-116: SetDevice, DeviceAlloc, InitCurand, then
// per complex SetParaT, CopyH2D, Dock; DeviceFree at the end) run against
// libgeauxdock_amd.so through include/geauxdock_dropin.h, on a complex filled
// into the reference's own Complex layout from the library's loaders.
//
//   dropin_main FIXTURE_DIR PARAM_DIR STEPS [EXCHANGE_PERIOD N_TMP]
//
// FIXTURE_DIR holds 1a07C1.sdf, 1a07C1-0.8.ff and 1a07C-1.pdb, PARAM_DIR the
// force field (paras, 08ff_opt, 08_nor_a, 08_nor_b).  Prints Dock's report (replica 0's
// first record of the last dump, the Ref Result line, timers, Benchmark CSV)
// and then one "final" line per replica: temperature index, accept flag and
// the ten terms of its final state.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/geauxdock_amd.h"
#include "../../include/geauxdock_dropin.h"

static void check(gd_status st, const char *what) {
  if (st != GD_OK) {
    std::fprintf(stderr, "%s: status %d: %s\n", what, (int)st, gd_last_error());
    std::exit(2);
  }
}

// The loaders' job in the reference Complex layout, as server1.C:54-217 and
// util_optimize_soa.C leave it: ligand file order in x2/t2/c2, type-sorted
// order in x3/t3/c3 (and x1 = x3), MCS columns in file order.
static void fill(const gd_job_desc &d, const gd_enepara &e, Complex *c) {
  const int L = d.lig_natom, P = d.prt_npoint, K = d.kde_npoint, M = d.mcs_nrow;
  for (int i = 0; i < d.n_lig; ++i) {
    Ligand &g = c->lig[i];
    g.lig_natom = L;
    for (int k = 0; k < 3; ++k) g.center[k] = d.lig_center[k];
    for (int j = 0; j < L; ++j) {
      const size_t q = (size_t)i * L + j;
      const int f = d.lig_perm ? d.lig_perm[j] : j;
      g.x1[j] = g.x3[j] = d.lig_x[q];
      g.y1[j] = g.y3[j] = d.lig_y[q];
      g.z1[j] = g.z3[j] = d.lig_z[q];
      g.t1[j] = g.t3[j] = d.lig_t[j];
      g.c1[j] = g.c3[j] = d.lig_c[j];
      g.x2[f] = d.lig_x[q], g.y2[f] = d.lig_y[q], g.z2[f] = d.lig_z[q];
      g.t2[f] = d.lig_t[j], g.c2[f] = d.lig_c[j];
    }
  }
  for (int i = 0; i < d.n_prt; ++i) {
    Protein &p = c->prt[i];
    p.prt_npoint = P;
    for (int k = 0; k < 3; ++k) p.pocket_center[k] = d.pocket_center[k];
    for (int j = 0; j < P; ++j) {
      const size_t q = (size_t)i * P + j;
      p.x[j] = d.prt_x[q], p.y[j] = d.prt_y[q], p.z[j] = d.prt_z[q];
      p.t[j] = d.prt_t[j], p.cdc[j] = d.prt_cdc[j], p.c[j] = d.prt_c[j];
      p.hpp[j] = d.prt_hpp[j], p.ele[j] = d.prt_ele[j], p.seq3r[j] = d.prt_seq3r[j];
    }
  }
  for (int lt = 0; lt < GD_NLT; ++lt)
    for (int s = 0; s < d.psp_ncol && s < GD_REF_MAXPRO; ++s)
      c->psp.psp[lt][s] = d.psp[(size_t)lt * d.psp_ncol + s];
  c->kde.kde_npoint = K;
  for (int k = 0; k < K; ++k)
    c->kde.t[k] = d.kde_t[k], c->kde.x[k] = d.kde_x[k], c->kde.y[k] = d.kde_y[k], c->kde.z[k] = d.kde_z[k];
  for (int m = 0; m < M; ++m) {
    Mcs &r = c->mcs[m];
    for (int j = 0; j < GD_REF_MAX_MCS_COL; ++j) r.x[j] = r.y[j] = r.z[j] = GD_MCS_INVALID;
    int n = 0;
    for (int j = 0; j < L; ++j) {
      const int f = d.lig_perm ? d.lig_perm[j] : j;
      const size_t q = (size_t)m * L + j;
      r.x[f] = d.mcs_x[q], r.y[f] = d.mcs_y[q], r.z[f] = d.mcs_z[q];
      n += d.mcs_x[q] != GD_MCS_INVALID;
    }
    r.ncol = n;
    r.tcc = d.mcs_tcc[m];
  }
  EnePara &p = c->enepara;
  std::memcpy(p.p1a, e.p1a, sizeof p.p1a);
  std::memcpy(p.p2a, e.p2a, sizeof p.p2a);
  std::memcpy(p.pmf0, e.pmf0, sizeof p.pmf0);
  std::memcpy(p.pmf1, e.pmf1, sizeof p.pmf1);
  std::memcpy(p.hdb0, e.hdb0, sizeof p.hdb0);
  std::memcpy(p.hdb1, e.hdb1, sizeof p.hdb1);
  p.lj0 = e.lj0, p.lj1 = e.lj1, p.el1 = e.el1, p.el0 = e.el0, p.a1 = e.a1, p.b1 = e.b1;
  std::memcpy(p.hpp, e.hpp, sizeof p.hpp);
  std::memcpy(p.hpl0, e.hpl0, sizeof p.hpl0);
  std::memcpy(p.hpl1, e.hpl1, sizeof p.hpl1);
  std::memcpy(p.hpl2, e.hpl2, sizeof p.hpl2);
  p.kde2 = e.kde2, p.kde3 = e.kde3;
  std::memcpy(p.a_para, e.a_para, sizeof p.a_para);
  std::memcpy(p.b_para, e.b_para, sizeof p.b_para);
  std::memcpy(p.w, e.w, sizeof p.w);
  const int nl = d.n_lig, np = d.n_prt, nt = d.n_tmp;
  for (int t = 0; t < nt; ++t) {
    c->temp[t].minus_beta = d.minus_beta[t];
    c->temp[t].t = -1.0f / d.minus_beta[t];
    c->temp[t].order = t;
  }
  for (int pp = 0; pp < np; ++pp)  // SetReplica, util.C:370-406
    for (int t = 0; t < nt; ++t)
      for (int l = 0; l < nl; ++l) {
        ReplicaMC &r = c->replica[nt * nl * pp + nl * t + l];
        r.idx_rep = nt * nl * pp + nl * t + l;
        r.idx_prt = pp, r.idx_tmp = t, r.idx_lig = l;
      }
  c->size.n_lig = nl, c->size.n_prt = np, c->size.n_tmp = nt, c->size.n_rep = nl * np * nt;
  c->size.lig_natom = L, c->size.prt_npoint = P, c->size.kde_npoint = K, c->size.mcs_nrow = M;
  std::strcpy(c->lig_file, "1a07C1");
  std::strcpy(c->prt_file, "1a07C-1");
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s FIXTURE_DIR PARAM_DIR STEPS [EXCHANGE_PERIOD N_TMP]\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1], par = argv[2];
  const int steps = std::atoi(argv[3]);
  const int spe = argc > 4 ? std::atoi(argv[4]) : 0;
  const int n_tmp = argc > 5 ? std::atoi(argv[5]) : 1;
  gd_enepara ene;
  check(gd_load_enepara((par + "/paras").c_str(), (par + "/08ff_opt").c_str(),
                        (par + "/08_nor_a").c_str(), (par + "/08_nor_b").c_str(), &ene),
        "gd_load_enepara");
  gd_host_job *h = nullptr;
  check(gd_load_complex((dir + "/1a07C1.sdf").c_str(), (dir + "/1a07C1-0.8.ff").c_str(),
                        (dir + "/1a07C-1.pdb").c_str(), "1a07C1", &ene, n_tmp, 0.044f,
                        n_tmp > 1 ? 0.4f : 0.036f, &h),
        "gd_load_complex");
  gd_job_desc d;
  check(gd_host_job_desc(h, &d), "gd_host_job_desc");

  // host data (main_gpu allocates these pinned; plain heap memory is enough here)
  Complex *ch = new Complex();
  std::vector<Record> rh(GD_REF_MAX_REP);
  fill(d, ene, ch);
  ch->mcpara.steps_total = steps;
  ch->mcpara.steps_per_dump = steps < GD_REF_STEPS_PER_DUMP ? (steps > 0 ? steps : 1) : GD_REF_STEPS_PER_DUMP;
  // the reference frontend never writes this field (McPara is allocated
  // uninitialised): leave a leftover value in it; the exchange period is given
  // through GD_EXCHANGE_STEPS instead
  ch->mcpara.steps_per_exchange = 3;
  if (spe > 0) setenv("GD_EXCHANGE_STEPS", std::to_string(spe).c_str(), 1);
  const float scale[6] = {0.02f, 0.02f, 0.02f, 0.08f, 0.08f, 0.08f};  // the reference defaults
  for (int k = 0; k < 6; ++k) ch->mcpara.move_scale[k] = scale[k];

  Complex *cd[1];
  Record *rd[1];
  gd_rng_state *sd[1];
  ParaT parat[1];
  ParaT *pt[1] = {&parat[0]};
  SetDevice();
  DeviceAlloc(cd, rd, sd);
  InitCurand(sd);
  SetParaT(ch, pt);
  CopyH2D(ch, cd, pt);
  Dock(ch, rh.data(), cd, rd, pt, sd);
  std::printf("sizes %d %d %d\n", ch->size.n_lig, ch->size.n_prt, ch->size.n_tmp);
  for (int r = 0; r < ch->size.n_rep; ++r) {
    const ReplicaMC &q = ch->replica[r];
    std::printf("final %d %d %d", r, q.idx_tmp, q.is_accept);
    for (int i = 0; i < 10; ++i) std::printf(" %.6f", q.energy[i]);
    std::printf(" records %d\n", rh[r].next_entry);
  }
  DeviceFree(cd, rd, sd);
  delete ch;
  gd_host_job_free(h);
  return 0;
}
