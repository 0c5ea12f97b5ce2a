// gd_oracle.cpp -- CPU oracle for the GeauxDock Monte Carlo hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
// check of __graft_entry__.py and the cpu_baseline leg of bench.py may load
// it; the product (geauxdock_amd/) never does.
//
// A restatement, in plain C++, of the reference CPU backend:
//   MonteCarlo_d            src/backend_cpu_mic/kernel_cpp_montecarlo_soa.C:32-418
//     move                  :242-250      rotation/transform  :256-293
//     e_s assembly          :318-324      DST                 :348-359
//     normalise + weight    :363-374      Metropolis accept   :382-392
//   PRT pair terms          src/backend_cpu_mic/energy_prt_v1.C:5-134
//   KDE term                src/backend_cpu_mic/energy_kde_v1.C:1-51
//   MCS term                src/backend_cpu_mic/energy_mcs_v1.C:1-31
//   Dock driver             src/backend_cpu_mic/backend_cpu_mic.C:34-98
// with the time-seeded MyRand_d (kernel_cpp_montecarlo_soa.C:8-15) replaced by
// Philox4x32-10 (own restatement below, pinned by the Random123 KATs of
// src/module/Random123/examples/kat_vectors:27-29) and with temperature
// exchange, which the reference declares but never implements
// (geauxdock.h:340-369), specified as: every steps_per_exchange steps, in round
// k, pair temperatures (t, t+1) with t = k mod 2, t+2, ... inside every
// (ligand conformer, protein conformer) ladder and swap their temperature
// labels iff u < exp((mb[t+1] - mb[t]) * (E_i - E_j)); then, per enabled
// conformer axis, pose swaps between neighbour conformers (swap_axis below).
//
// Arithmetic modes:
//   FP32 "gpu" rounding  : the squared distance is fmaf(dx,dx,fmaf(dy,dy,dz*dz)),
//                          MCS sums run in type-sorted atom order and the DST
//                          offset is m + (center - pocket) -- one fixed rounding
//                          sequence shared with the HIP kernels, so the
//                          discontinuous psp cutoff decides identically on
//                          identical coordinates;
//   FP32 "literal"       : the reference's own expression order (dx*dx+dy*dy+
//                          dz*dz, MCS in file atom order); with -ffp-contract=off
//                          it reproduces a g++ build of the reference bit for bit;
//   FP64 twin            : the same algorithm in double, the error reference.
// Compiled with -ffp-contract=off so no fused operations appear implicitly.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "../include/geauxdock_amd.h"

namespace {

// ---------------------------------------------------------------- Philox
struct P4 {
  uint32_t v[4];
};
P4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, int R) {
  for (int r = 0; r < R; ++r) {
    const uint64_t a = (uint64_t)0xD2511F53u * c0, b = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t y0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
    const uint32_t y1 = (uint32_t)b;
    const uint32_t y2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
    const uint32_t y3 = (uint32_t)a;
    c0 = y0, c1 = y1, c2 = y2, c3 = y3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return P4{{c0, c1, c2, c3}};
}
inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

struct Draws {
  uint32_t w[8];
};
Draws draws(uint64_t seed, uint32_t job, uint32_t rep, uint32_t step) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const P4 a = philox(step, rep, job, 0, k0, k1, 10), b = philox(step, rep, job, 1, k0, k1, 10);
  Draws d;
  for (int i = 0; i < 4; ++i) d.w[i] = a.v[i], d.w[4 + i] = b.v[i];
  return d;
}

// move increments for one replica-step (reference: scale * U[0,1))
void move_incr(const gd_mc_param *mc, const Draws &d, float inc[6]) {
  if (mc->move_mode == GD_MOVE_GAUSSIAN) {
    for (int k = 0; k < 3; ++k) {
      const float u1 = u01(d.w[2 * k]), u2 = u01(d.w[2 * k + 1]);
      const float rad = sqrtf(-2.0f * logf(1.0f - u1));
      const float th = 6.28318530717958647692f * u2;
      inc[2 * k] = mc->move_scale[2 * k] * (rad * cosf(th));
      inc[2 * k + 1] = mc->move_scale[2 * k + 1] * (rad * sinf(th));
    }
  } else {
    for (int i = 0; i < 6; ++i) inc[i] = mc->move_scale[i] * u01(d.w[i]);
  }
}

// ------------------------------------------------------------ energy
struct Job {
  const gd_job_desc *d;
  int L, P, K, M;
  std::vector<int> file_of;  // sorted atom j -> file index
  std::vector<int> at_file;  // file index -> sorted atom j
};

Job make_job(const gd_job_desc *d) {
  Job j;
  j.d = d;
  j.L = d->lig_natom, j.P = d->prt_npoint, j.K = d->kde_npoint, j.M = d->mcs_nrow;
  j.file_of.resize(j.L);
  j.at_file.resize(j.L);
  for (int a = 0; a < j.L; ++a) j.file_of[a] = d->lig_perm ? d->lig_perm[a] : a;
  for (int a = 0; a < j.L; ++a) j.at_file[j.file_of[a]] = a;
  return j;
}

template <typename T>
struct Out {
  T raw[9];    // vdw ele pmf psp hdb*(-1/sqrt(2pi)) hpc kde*(1/kde3) lhm dst
  T term[10];  // normalised terms and total
  double scale[9];  // sum of |contributions|, the error scale of each raw term
};

template <typename T>
inline T t_sqrt(T x);
template <>
inline float t_sqrt(float x) { return sqrtf(x); }
template <>
inline double t_sqrt(double x) { return sqrt(x); }
template <typename T>
inline T t_exp(T x);
template <>
inline float t_exp(float x) { return expf(x); }
template <>
inline double t_exp(double x) { return exp(x); }
template <typename T>
inline T t_log(T x);
template <>
inline float t_log(float x) { return logf(x); }
template <>
inline double t_log(double x) { return log(x); }

// Ligand transform of kernel_cpp_montecarlo_soa.C:256-293 (float in the
// FP32 modes, double in the FP64 twin).
template <typename T>
void transform(const gd_job_desc *d, int lig, const float mf[6], T *x, T *y, T *z) {
  const int L = d->lig_natom;
  T m[6];
  for (int i = 0; i < 6; ++i) m[i] = (T)mf[i];
  T s1, c1, s2, c2, s3, c3;
  if (sizeof(T) == 4) {
    s1 = sinf((float)m[3]), c1 = cosf((float)m[3]);
    s2 = sinf((float)m[4]), c2 = cosf((float)m[4]);
    s3 = sinf((float)m[5]), c3 = cosf((float)m[5]);
  } else {
    s1 = sin((double)m[3]), c1 = cos((double)m[3]);
    s2 = sin((double)m[4]), c2 = cos((double)m[4]);
    s3 = sin((double)m[5]), c3 = cos((double)m[5]);
  }
  T r[3][3];
  r[0][0] = c1 * c2;
  r[0][1] = c1 * s2 * s3 - c3 * s1;
  r[0][2] = s1 * s3 + c1 * c3 * s2;
  r[1][0] = c2 * s1;
  r[1][1] = c1 * c3 + s1 * s2 * s3;
  r[1][2] = c3 * s1 * s2 - c1 * s3;
  r[2][0] = -1 * s2;
  r[2][1] = c2 * s3;
  r[2][2] = c2 * c3;
  const float *cx = d->lig_x + (size_t)lig * L, *cy = d->lig_y + (size_t)lig * L,
              *cz = d->lig_z + (size_t)lig * L;
  for (int l = 0; l < L; ++l) {
    const T X = cx[l], Y = cy[l], Z = cz[l];
    x[l] = r[0][0] * X + r[0][1] * Y + r[0][2] * Z + m[0] + (T)d->lig_center[0];
    y[l] = r[1][0] * X + r[1][1] * Y + r[1][2] * Z + m[1] + (T)d->lig_center[1];
    z[l] = r[2][0] * X + r[2][1] * Y + r[2][2] * Z + m[2] + (T)d->lig_center[2];
  }
}

// Full energy at given transformed coordinates (x,y,z in type-sorted order).
// literal = reference expression order; otherwise the shared fma rounding.
template <typename T>
void energy(const Job &J, int prt, const float m[6], const T *xf, const T *yf, const T *zf,
            bool literal, Out<T> &o) {
  const gd_job_desc *d = J.d;
  const gd_enepara *e = d->ene;
  const int L = J.L, P = J.P, K = J.K, M = J.M;
  const float *px = d->prt_x + (size_t)prt * P, *py = d->prt_y + (size_t)prt * P,
              *pz = d->prt_z + (size_t)prt * P;
  T evdw = 0, eele = 0, epmf = 0, epsp = 0, ehdb = 0, ehpc = 0;
  double sc[9] = {0};
  const T lj0 = e->lj0, lj1 = e->lj1, el0 = e->el0, el1 = e->el1, a1 = e->a1, b1 = e->b1;
  for (int l = 0; l < L; ++l) {  // energy_prt_v1.C:5
    T ehpc1 = 0;
    double shpc1 = 0;
    const int lt = d->lig_t[l];
    const T lx = xf[l], ly = yf[l], lz = zf[l], lq = d->lig_c[l];
    for (int p = 0; p < P; ++p) {  // :30
      const int pt = d->prt_t[p];
      const T dx = lx - (T)px[p], dy = ly - (T)py[p], dz = lz - (T)pz[p];
      T d2;
      if (literal || sizeof(T) == 8)
        d2 = dx * dx + dy * dy + dz * dz;
      else
        d2 = (T)fmaf((float)dx, (float)dx, fmaf((float)dy, (float)dy, (float)(dz * dz)));
      const T d4 = d2 * d2;
      const T dst = t_sqrt(d2);
      if (d->prt_cdc[p] == 1 && d2 <= (T)81.0f) {  // :44-50
        const T h = (T)d->prt_hpp[p] *
                    ((T)1.0f - ((T)(3.5f / 81.0f) * d2 - (T)(4.5f / 81.0f / 81.0f) * d4 +
                                (T)(2.5f / 81.0f / 81.0f / 81.0f) * d4 * d2 -
                                (T)(0.5f / 81.0f / 81.0f / 81.0f / 81.0f) * d4 * d4));
        ehpc1 += h;
        shpc1 += fabs((double)h);
      }
      {  // L-J, :61-64
        const T p1 = (T)e->p1a[lt][pt] / (d4 * d4 * dst);
        const T p2 = (T)e->p2a[lt][pt] / (d4 * d2);
        const T p4 = p1 * lj0 * ((T)1.0f + lj1 * d2) + (T)1.0f;
        const T v = (p1 - p2) / p4;
        evdw += v;
        sc[0] += fabs((double)v);
      }
      {  // electrostatics, :72-78
        const T s1 = el1 * dst;
        T g1;
        if (s1 < (T)1)
          g1 = el0 + a1 * s1 * s1 + b1 * s1 * s1 * s1;
        else
          g1 = (T)1.0f / s1;
        const T v = lq * (T)d->prt_ele[p] * g1;
        eele += v;
        sc[1] += fabs((double)v);
      }
      const T dmp = dst - (T)e->pmf0[lt][pt];
      {  // contact, :88-91
        const T v = (T)e->pmf1[lt][pt] / ((T)1.0f + t_exp(((T)-0.5f * dst + (T)6.0f) * dmp));
        epmf += v;
        sc[2] += fabs((double)v);
      }
      if (d->prt_c[p] == 2 && dmp <= (T)0) {  // pocket-specific, :100-107
        const int res = d->prt_seq3r[p];
        const T v = (res >= 0 && res < d->psp_ncol) ? (T)d->psp[(size_t)lt * d->psp_ncol + res] : (T)0;
        epsp += v;
        sc[3] += fabs((double)v);
      }
      const T hdb0 = e->hdb0[lt][pt];
      if (hdb0 > (T)0.1f) {  // hydrogen bond, :116-121
        const T hdb1 = e->hdb1[lt][pt];
        T v;
        if (literal || sizeof(T) == 8) {
          const T h3 = (dst - hdb0) * hdb1;
          v = hdb1 * t_exp((T)-0.5f * h3 * h3);
        } else {
          // GPU rounding: the exponent q = log2(e)/2 ((r - hdb0) hdb1)^2 in FP64
          // from exact coordinate differences, 2^-q = 2^-floor(q) 2^-frac(q)
          const double ex = (double)lx - (double)px[p], ey = (double)ly - (double)py[p],
                       ez = (double)lz - (double)pz[p];
          const double rd = sqrt(ex * ex + ey * ey + ez * ez);
          const double hd = (rd - (double)e->hdb0[lt][pt]) * (double)e->hdb1[lt][pt];
          const double q = hd * hd * 0.72134752044448170368;
          const double fl = floor(q);
          v = hdb1 * ldexpf(exp2f(-(float)(q - fl)), -(int)fl);
        }
        ehdb += v;
        sc[4] += fabs((double)v);
      }
    }
    const T hpc2 = (ehpc1 - (T)e->hpl0[lt]) / (T)e->hpl1[lt];  // :132-133
    const T hv = (T)0.5f * hpc2 * hpc2 - (T)e->hpl2[lt];
    ehpc += hv;
    // error scale: |term| + sensitivity to the per-atom sum
    sc[5] += fabs((double)hv) + fabs((double)hpc2 / (double)e->hpl1[lt]) * shpc1;
  }
  const T sqrt_2_pi_inv = (T)(-1.0f / sqrtf(2.0f * 3.1415926535f));
  T es[10];
  es[0] = evdw;
  es[1] = eele;
  es[2] = epmf;
  es[3] = epsp;
  es[4] = ehdb * sqrt_2_pi_inv;
  sc[4] *= fabs((double)sqrt_2_pi_inv);
  es[5] = ehpc;
  {  // KDE, energy_kde_v1.C
    T ekde = 0;
    for (int l = 0; l < L; ++l) {
      T s = 0;
      int n = 0;
      const int lt = d->lig_t[l];
      for (int k = 0; k < K; ++k) {
        if (lt == d->kde_t[k]) {
          const T dx = (T)xf[l] - (T)d->kde_x[k], dy = (T)yf[l] - (T)d->kde_y[k],
                  dz = (T)zf[l] - (T)d->kde_z[k];
          const T d2 = dx * dx + dy * dy + dz * dz;
          s += t_exp((T)e->kde2 * d2);
          ++n;
        }
      }
      if (n != 0) ekde += s / (T)n;
    }
    const T kde3_inv = (T)(1.0f / e->kde3);
    es[6] = ekde * kde3_inv;
    sc[6] = fabs((double)es[6]);
  }
  {  // MCS, energy_mcs_v1.C
    T elhm = 0;
    double s_abs = 0;
    for (int m = 0; m < M; ++m) {
      T e1 = 0;
      int n = 0;
      for (int q = 0; q < L; ++q) {
        const int a = literal ? J.at_file[q] : q;  // literal: file atom order
        const size_t idx = (size_t)m * L + a;
        if (d->mcs_x[idx] != GD_MCS_INVALID) {
          const T dx = (T)xf[a] - (T)d->mcs_x[idx], dy = (T)yf[a] - (T)d->mcs_y[idx],
                  dz = (T)zf[a] - (T)d->mcs_z[idx];
          e1 += dx * dx + dy * dy + dz * dz;
          ++n;
        }
      }
      const T v = (T)d->mcs_tcc[m] * t_sqrt(e1 / (T)n);
      elhm += v;
      s_abs += fabs((double)v);
    }
    es[7] = t_log(elhm / (T)M);
    sc[7] = std::max(1.0, fabs((double)es[7]));
    (void)s_abs;
  }
  {  // DST, kernel_cpp_montecarlo_soa.C:348-359; the literal mode keeps the
     // reference's order, the other modes compute m + (center - pocket) with the
     // difference rounded once, as the kernel does
    T dst = 0;
    for (int i = 0; i < 3; ++i) {
      const T dd = literal ? (T)d->lig_center[i] + (T)m[i] - (T)d->pocket_center[i]
                           : (T)m[i] + (T)((double)d->lig_center[i] - (double)d->pocket_center[i]);
      dst += dd * dd;
    }
    es[8] = t_sqrt(dst);
    sc[8] = fabs((double)es[8]);
  }
  for (int i = 0; i < 9; ++i) o.raw[i] = es[i];
  for (int i = 0; i < 9; ++i) o.scale[i] = sc[i] > 0 ? sc[i] : fabs((double)es[i]);
  // normalisation and weighting, :363-374
  for (int i = 0; i < 7; ++i) es[i] = es[i] / (T)L;
  for (int i = 0; i < 9; ++i) es[i] = (T)e->a_para[i] * es[i] + (T)e->b_para[i];
  T et = 0;
  for (int i = 0; i < 9; ++i) et += (T)e->w[i] * es[i];
  es[9] = et;
  for (int i = 0; i < 10; ++i) o.term[i] = es[i];
}

// Metropolis test of :382-385; the forced initial accept draws u = 0.
inline bool accept(float u, float de, float mb) { return u < expf(de * mb); }

// Temperature exchange of one ladder in round `round`: pairs (t, t+1) for
// t = round mod 2, t+2, ...; swap temperature labels iff
// u < exp((mb[t+1] - mb[t]) * (E_i - E_j)), u from counter
// (round, ladder*n_tmp + t, job, 2).
void exchange_ladder(const gd_mc_param *mc, const float *mb, int nt, uint32_t ladder, int round,
                     int *own, const float *en, int *tmp, int *nexc) {
  for (int t = round & 1; t + 1 < nt; t += 2) {
    const int i = own[t], j = own[t + 1];
    const P4 w = philox((uint32_t)round, ladder * (uint32_t)nt + (uint32_t)t, mc->job_id, 2u,
                        (uint32_t)mc->seed, (uint32_t)(mc->seed >> 32), 10);
    const float u = u01(w.v[0]);
    const float x = (mb[t + 1] - mb[t]) * (en[(size_t)i * 10 + 9] - en[(size_t)j * 10 + 9]);
    if (u < expf(x)) {
      own[t] = j, own[t + 1] = i;
      tmp[i] = t + 1, tmp[j] = t;
      ++nexc[i], ++nexc[j];
    }
  }
}

// Partner of replica r for a pose swap along `axis` (0: ligand conformer,
// 1: protein conformer) in round `round`: the replicas that share the other
// conformer index and the temperature form a line; conformers (c, c+1) pair
// for c = round mod 2, c+2, ...  Returns -1 when r is unpaired.
int swap_partner(int r, int axis, int round, int nl, int np, int nt, const int *lig, const int *prt,
                 const int *tmp, const int *owner, bool *lower) {
  const int l = lig[r], p = prt[r], t = tmp[r];
  const int n = axis ? np : nl, c = axis ? p : l;
  *lower = false;
  if (c - (round & 1) < 0) return -1;
  const int cq = ((c - (round & 1)) & 1) ? c - 1 : c + 1;
  if (cq >= n) return -1;
  *lower = c < cq;
  const int p2 = axis ? cq : p, l2 = axis ? l : cq;
  return owner[((size_t)p2 * nl + l2) * nt + t];
}

// Joint Metropolis decision of a pose-swap pair (a: lower-conformer member,
// b: partner; E the totals before, T the totals at the partner's pose):
// accept iff u < exp(mb_a (T_a - E_a) + mb_b (T_b - E_b)), u from Philox
// counter (round, a, job, 3 + axis).
bool swap_accept(const gd_mc_param *mc, int axis, int round, int a, float mba, float ta, float ea,
                 float mbb, float tb, float eb) {
  const P4 w = philox((uint32_t)round, (uint32_t)a, mc->job_id, 3u + (uint32_t)axis,
                      (uint32_t)mc->seed, (uint32_t)(mc->seed >> 32), 10);
  return u01(w.v[0]) < expf(mba * (ta - ea) + mbb * (tb - eb));
}

}  // namespace

extern "C" {

// Pose-swap rule on a frozen two-conformer toy: conformer 0 starts at pose X,
// conformer 1 at pose Y; e[c][x] is conformer c's total at pose x (x: 0 = X,
// 1 = Y).  Runs `rounds` attempts of the pair (replica 0 on conformer 0,
// replica 1 on conformer 1, inverse temperature -mb for both) and returns the
// swaps per attempt.
double or_swap_toy(const float e[4], float mb, int rounds, uint64_t seed) {
  gd_mc_param mc;
  memset(&mc, 0, sizeof mc);
  mc.seed = seed;
  int x0 = 0;  // pose held by conformer 0; conformer 1 holds 1 - x0
  long swaps = 0;
  for (int k = 0; k < rounds; ++k) {
    const int x1 = 1 - x0;
    if (swap_accept(&mc, 0, k, 0, mb, e[0 * 2 + x1], e[0 * 2 + x0], mb, e[1 * 2 + x0], e[1 * 2 + x1])) {
      x0 = x1;
      ++swaps;
    }
  }
  return rounds ? (double)swaps / rounds : 0.0;
}

// Exchange rule on a frozen ladder with the given total energies: runs
// `rounds` exchange rounds, returns swaps per attempted pair and the final
// temperature of each replica (test of the acceptance rule in isolation).
double or_exchange_toy(int nt, const float *energy, const float *mb, int rounds, uint64_t seed,
                       int *final_tmp) {
  gd_mc_param mc;
  memset(&mc, 0, sizeof mc);
  mc.seed = seed;
  std::vector<float> en((size_t)nt * 10, 0.0f);
  std::vector<int> own(nt), tmp(nt), nexc(nt, 0);
  for (int t = 0; t < nt; ++t) en[(size_t)t * 10 + 9] = energy[t], own[t] = t, tmp[t] = t;
  long attempts = 0;
  for (int k = 0; k < rounds; ++k) {
    for (int t = k & 1; t + 1 < nt; t += 2) ++attempts;
    exchange_ladder(&mc, mb, nt, 0u, k, own.data(), en.data(), tmp.data(), nexc.data());
  }
  long swaps = 0;
  for (int t = 0; t < nt; ++t) swaps += nexc[t];
  if (final_tmp)
    for (int t = 0; t < nt; ++t) final_tmp[t] = tmp[t];
  return attempts ? (double)swaps / 2.0 / (double)attempts : 0.0;
}


// Philox4x32-R for the known-answer tests.
void or_philox(const uint32_t ctr[4], const uint32_t key[2], int rounds, uint32_t out[4]) {
  const P4 r = philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], rounds);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

// Move increments of the oracle's chains: inc[s][r][6] for replicas r < n_rep
// at steps step0 + s (the statistical tests of the move generator).
void or_move_draws(uint64_t seed, uint32_t job, int move_mode, const float scale[6], int n_rep,
                   uint32_t step0, int n_steps, float *inc) {
  gd_mc_param mc;
  memset(&mc, 0, sizeof mc);
  mc.move_mode = move_mode;
  for (int k = 0; k < 6; ++k) mc.move_scale[k] = scale[k];
  for (long long i = 0; i < (long long)n_rep * n_steps; ++i) {
    const Draws d = draws(seed, job, (uint32_t)(i % n_rep), step0 + (uint32_t)(i / n_rep));
    move_incr(&mc, d, inc + i * 6);
  }
}

void or_step_draws(uint64_t seed, uint32_t job, uint32_t rep, uint32_t step, uint32_t w[8]) {
  const Draws d = draws(seed, job, rep, step);
  for (int i = 0; i < 8; ++i) w[i] = d.w[i];
}

// Energies of poses.  mode: 0 = FP32 shared rounding, 1 = FP32 literal,
// 2 = FP64.  coords_in (optional) replaces the transform: [n][L][3].
int or_energy(const gd_job_desc *d, int n, const int *conf, const float *pose,
              const float *coords_in, int mode, double *terms, double *raw, double *scale,
              float *coords_out, int nthreads) {
  if (!d || n < 0 || !conf || !pose) return GD_ERR_ARG;
  const Job J = make_job(d);
  const int L = J.L;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n; ++i) {
    const float *m = pose + 6 * (size_t)i;
    if (mode == 2) {
      std::vector<double> x(L), y(L), z(L);
      if (coords_in) {
        for (int l = 0; l < L; ++l) {
          x[l] = coords_in[((size_t)i * L + l) * 3 + 0];
          y[l] = coords_in[((size_t)i * L + l) * 3 + 1];
          z[l] = coords_in[((size_t)i * L + l) * 3 + 2];
        }
      } else {
        transform<double>(d, conf[2 * i], m, x.data(), y.data(), z.data());
      }
      if (coords_out)
        for (int l = 0; l < L; ++l) {
          coords_out[((size_t)i * L + l) * 3 + 0] = (float)x[l];
          coords_out[((size_t)i * L + l) * 3 + 1] = (float)y[l];
          coords_out[((size_t)i * L + l) * 3 + 2] = (float)z[l];
        }
      Out<double> o;
      energy<double>(J, conf[2 * i + 1], m, x.data(), y.data(), z.data(), false, o);
      for (int k = 0; k < 10; ++k)
        if (terms) terms[(size_t)i * 10 + k] = o.term[k];
      for (int k = 0; k < 9; ++k) {
        if (raw) raw[(size_t)i * 9 + k] = o.raw[k];
        if (scale) scale[(size_t)i * 9 + k] = o.scale[k];
      }
    } else {
      std::vector<float> x(L), y(L), z(L);
      if (coords_in) {
        for (int l = 0; l < L; ++l) {
          x[l] = coords_in[((size_t)i * L + l) * 3 + 0];
          y[l] = coords_in[((size_t)i * L + l) * 3 + 1];
          z[l] = coords_in[((size_t)i * L + l) * 3 + 2];
        }
      } else {
        transform<float>(d, conf[2 * i], m, x.data(), y.data(), z.data());
      }
      if (coords_out)
        for (int l = 0; l < L; ++l) {
          coords_out[((size_t)i * L + l) * 3 + 0] = x[l];
          coords_out[((size_t)i * L + l) * 3 + 1] = y[l];
          coords_out[((size_t)i * L + l) * 3 + 2] = z[l];
        }
      Out<float> o;
      energy<float>(J, conf[2 * i + 1], m, x.data(), y.data(), z.data(), mode == 1, o);
      for (int k = 0; k < 10; ++k)
        if (terms) terms[(size_t)i * 10 + k] = o.term[k];
      for (int k = 0; k < 9; ++k) {
        if (raw) raw[(size_t)i * 9 + k] = o.raw[k];
        if (scale) scale[(size_t)i * 9 + k] = o.scale[k];
      }
    }
  }
  return GD_OK;
}

// Monte Carlo docking of one complex (FP32 shared rounding): initial-energy
// pass with forced accept, steps_total steps, an exchange point every
// steps_per_exchange steps (temperature exchange and conformer pose swaps as
// mc->exchange_axes selects).  Replicas in SetReplica order (util.C:370-406).
// rep_lo/rep_hi restrict the run to a replica range (for bounded timing
// samples; exchange then needs the whole range of each ladder).  Optional
// trace[n_rep][steps_total] receives the accepted total energy after each step.
// Optional rec[n_rep][rec_cap][17] / rec_count[n_rep] receive the accept
// records of the whole run (kernel_cpp_montecarlo_soa.C:224-231, never reset):
// at the start of step s, the replica's accepted (step s, movematrix,
// energy) when the step before it, the initial pass or a conformer swap was
// accepted.
// literal != 0 evaluates energies in the reference's own FP32 expression order
// (bit-exact against the reference backend) instead of the GPU rounding.
int or_dock(const gd_job_desc *d, const gd_mc_param *mc, int nthreads, int rep_lo, int rep_hi,
            gd_replica_out *out, float *trace, double *ms, float *rec, int rec_cap,
            int *rec_count, int literal) {
  if (!d || !mc) return GD_ERR_ARG;
  const Job J = make_job(d);
  const int L = J.L;
  const int nl = d->n_lig, np = d->n_prt, nt = d->n_tmp;
  const int R = nl * np * nt;
  if (rep_hi <= 0 || rep_hi > R) rep_hi = R;
  if (rep_lo < 0) rep_lo = 0;
  std::vector<float> mm((size_t)R * 6, 0.0f), en((size_t)R * 10, 0.0f), best(R, 1e30f);
  std::vector<int> tmp(R), lig(R), prt(R), acc(R, 0), nacc(R, 0), nexc(R, 0), nswp((size_t)2 * R, 0);
  const int axes = mc->exchange_axes ? mc->exchange_axes : GD_EXCH_TEMP;
  std::vector<int> owner((size_t)nl * np * nt);
  for (int p = 0; p < np; ++p)
    for (int t = 0; t < nt; ++t)
      for (int l = 0; l < nl; ++l) {
        const int r = nt * nl * p + nl * t + l;  // util.C:380
        lig[r] = l, prt[r] = p, tmp[r] = t;
        owner[((size_t)p * nl + l) * nt + t] = r;
      }
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (rec_count)
    for (int r = 0; r < R; ++r) rec_count[r] = 0;
  if (!rec_count) rec = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  const int seg = mc->steps_per_exchange > 0 ? mc->steps_per_exchange : std::max(1, mc->steps_total);
  // initial energy, forced accept (backend_cpu_mic.C:55)
  for (int s0 = -1; s0 < mc->steps_total; s0 += (s0 < 0 ? 1 : seg)) {
    const int s_begin = s0 < 0 ? 0 : s0;
    const int s_end = s0 < 0 ? 0 : std::min(mc->steps_total, s0 + seg);
    const bool init = s0 < 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = rep_lo; r < rep_hi; ++r) {
      std::vector<float> x(L), y(L), z(L);
      Out<float> o;
      const int nsteps = init ? 1 : s_end - s_begin;
      for (int k = 0; k < nsteps; ++k) {
        const int s = s_begin + k;
        if (!init && rec && acc[r] && rec_count[r] < rec_cap) {
          float *o = rec + ((size_t)r * rec_cap + rec_count[r]) * 17;
          o[0] = (float)s;
          for (int i = 0; i < 6; ++i) o[1 + i] = mm[(size_t)r * 6 + i];
          for (int i = 0; i < 10; ++i) o[7 + i] = en[(size_t)r * 10 + i];
          ++rec_count[r];
        }
        float m[6], u = 0.0f;
        if (init) {
          for (int i = 0; i < 6; ++i) m[i] = mm[(size_t)r * 6 + i];
        } else {
          const Draws dr = draws(mc->seed, mc->job_id, (uint32_t)r, (uint32_t)s);
          float inc[6];
          move_incr(mc, dr, inc);
          for (int i = 0; i < 6; ++i) m[i] = inc[i] + mm[(size_t)r * 6 + i];
          u = u01(dr.w[6]);
        }
        transform<float>(d, lig[r], m, x.data(), y.data(), z.data());
        energy<float>(J, prt[r], m, x.data(), y.data(), z.data(), literal != 0, o);
        const float de = o.term[9] - en[(size_t)r * 10 + 9];
        const bool a = accept(u, de, d->minus_beta[tmp[r]]);
        acc[r] = a;
        if (a) {
          for (int i = 0; i < 10; ++i) en[(size_t)r * 10 + i] = o.term[i];
          for (int i = 0; i < 6; ++i) mm[(size_t)r * 6 + i] = m[i];
          if (!init) ++nacc[r];
          best[r] = std::min(best[r], (float)o.term[9]);
        }
        if (trace && !init) trace[(size_t)r * mc->steps_total + s] = en[(size_t)r * 10 + 9];
      }
    }
    // exchange point after each full segment: temperatures, then conformer swaps
    if (init || mc->steps_per_exchange <= 0 || s_end - s_begin != seg) continue;
    const int round = s_end / seg - 1;
    if ((axes & GD_EXCH_TEMP) && nt > 1) {
      for (int p = 0; p < np; ++p)
        for (int l = 0; l < nl; ++l) {
          int *own = &owner[((size_t)p * nl + l) * nt];
          bool inside = true;
          for (int t = 0; t < nt; ++t) inside = inside && own[t] >= rep_lo && own[t] < rep_hi;
          if (!inside) continue;
          exchange_ladder(mc, d->minus_beta, nt, (uint32_t)((size_t)p * nl + l), round, own,
                          en.data(), tmp.data(), nexc.data());
        }
    }
    for (int axis = 0; axis < 2; ++axis) {
      if (!(axes & (GD_EXCH_LIG << axis)) || (axis ? np : nl) < 2) continue;
      // every member of a pair evaluates its own conformers at the partner's pose
      std::vector<int> part(R, -1);
      std::vector<char> low(R, 0);
      for (int r = rep_lo; r < rep_hi; ++r) {
        bool lower;
        const int q = swap_partner(r, axis, round, nl, np, nt, lig.data(), prt.data(), tmp.data(),
                                   owner.data(), &lower);
        if (q >= rep_lo && q < rep_hi) part[r] = q, low[r] = lower;
      }
      std::vector<float> tr((size_t)R * 10, 0.0f);
#pragma omp parallel for schedule(dynamic, 1)
      for (int r = rep_lo; r < rep_hi; ++r) {
        if (part[r] < 0) continue;
        std::vector<float> x(L), y(L), z(L);
        Out<float> o;
        const float *m = &mm[(size_t)part[r] * 6];
        transform<float>(d, lig[r], m, x.data(), y.data(), z.data());
        energy<float>(J, prt[r], m, x.data(), y.data(), z.data(), literal != 0, o);
        for (int i = 0; i < 10; ++i) tr[(size_t)r * 10 + i] = o.term[i];
      }
      // joint decisions from the pre-swap state, then the updates
      std::vector<char> take(R, 0);
      for (int r = rep_lo; r < rep_hi; ++r) {
        if (part[r] < 0) continue;
        const int a = low[r] ? r : part[r], b = low[r] ? part[r] : r;
        take[r] = swap_accept(mc, axis, round, a, d->minus_beta[tmp[a]], tr[(size_t)a * 10 + 9],
                              en[(size_t)a * 10 + 9], d->minus_beta[tmp[b]], tr[(size_t)b * 10 + 9],
                              en[(size_t)b * 10 + 9]);
      }
      const std::vector<float> mm0 = mm;
      for (int r = rep_lo; r < rep_hi; ++r) {
        if (!take[r]) continue;
        for (int i = 0; i < 10; ++i) en[(size_t)r * 10 + i] = tr[(size_t)r * 10 + i];
        for (int i = 0; i < 6; ++i) mm[(size_t)r * 6 + i] = mm0[(size_t)part[r] * 6 + i];
        best[r] = std::min(best[r], tr[(size_t)r * 10 + 9]);
        ++nswp[(size_t)axis * R + r];
        acc[r] = 1;
      }
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (out) {
    for (int r = 0; r < R; ++r) {
      if (out->movematrix)
        for (int i = 0; i < 6; ++i) out->movematrix[(size_t)r * 6 + i] = mm[(size_t)r * 6 + i];
      if (out->energy)
        for (int i = 0; i < 10; ++i) out->energy[(size_t)r * 10 + i] = en[(size_t)r * 10 + i];
      if (out->idx_tmp) out->idx_tmp[r] = tmp[r];
      if (out->n_accept) out->n_accept[r] = nacc[r];
      if (out->n_exchange) out->n_exchange[r] = nexc[r];
      if (out->best_energy) out->best_energy[r] = best[r];
      if (out->is_accept) out->is_accept[r] = acc[r];
      if (out->n_swap_lig) out->n_swap_lig[r] = nswp[r];
      if (out->n_swap_prt) out->n_swap_prt[r] = nswp[(size_t)R + r];
    }
  }
  return GD_OK;
}

// The CPU backend's entry point Dock(Complex*, Record*)
// (src/backend_cpu_mic/backend_cpu_mic.h:4) in C-ABI form: whole replica
// range, every host thread, the Philox stream of mc->seed / mc->job_id.
int gd_cpu_dock(const gd_job_desc *d, const gd_mc_param *mc, gd_replica_out *out) {
  if (!d) return GD_ERR_ARG;
  return or_dock(d, mc, 0, 0, d->n_lig * d->n_prt * d->n_tmp, out, nullptr, nullptr, nullptr, 0,
                 nullptr, 1);
}

// Smallest float x with expf(x) > 0: the forced initial accept of the
// reference fails below it (expf underflows to zero).
float or_exp_underflow_threshold(void) {
  float x = -80.0f;
  while (expf(nextafterf(x, -1000.0f)) > 0.0f) x = nextafterf(x, -1000.0f);
  return x;
}

}  // extern "C"
