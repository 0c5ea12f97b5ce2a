// gd_kernels.hip -- hand-written gfx950 kernels of the Monte Carlo hot path.
//
// One MC step of every replica is three launches:
//   gd_pair_kernel   one 8-wave work group per item (64 replicas x one
//                    ligand-atom chunk x eight protein tiles, one per wave):
//                    PRT pair terms, the tiles' slices of the KDE anchors and
//                    of the MCS rows; the group sums its tiles in LDS and
//                    writes partial sums.  Lanes are replicas, so protein
//                    attributes, force-field rows, KDE anchors and MCS targets
//                    are wave-uniform scalar loads and the inner loop is pure
//                    FP32 VALU on ligand coordinates held in VGPRs.
//   gd_reduce_kernel 16 work groups per 64-replica block: sum the partial
//                    sums and apply the non-linear per-atom (hpc) and per-row
//                    (MCS) terms; one slice of 8 values per group.
//   gd_finish_kernel one work group per 64-replica block: sums the slices,
//                    normalises and weights, records, runs the Metropolis
//                    test, and proposes + transforms the next move
//                    (Philox4x32-10, counter-based).
// plus gd_transform_kernel (pose -> ligand coordinates, used before the first
// pass and for energy evaluation), gd_exchange_kernel (temperature swaps) and
// the pose-swap kernels gd_swap_propose_kernel / gd_swap_accept_kernel.
//
// The reference computes the same quantities in one CUDA block per replica
// (backend_gpu/kernel_cuda_l1_montecarlo.cu:29-536); here the pair loop
// formulas follow backend_cpu_mic/energy_prt_v1.C:30-134, the KDE term
// energy_kde_v1.C:1-51, the MCS term energy_mcs_v1.C:1-31 and the
// step/accept logic kernel_cpp_montecarlo_soa.C:219-413.
#include <hip/hip_runtime.h>

#include "gd_device.h"
#include "gd_rng.h"

#define GD_LOG2E 1.44269504088896340736f
// Smallest float x with expf(x) > 0 in IEEE arithmetic with subnormals: the
// reference's forced initial accept (u = 0 < expf(dE * beta),
// kernel_cpp_montecarlo_soa.C:384-385) fails below it.
#define GD_EXP_UNDERFLOW (-103.97207641601562f)

__device__ __forceinline__ int gd_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
// Lane id within the wave, computed afresh (volatile: never reused from an
// earlier value, so a lane-derived index need not stay live across a loop).
__device__ __forceinline__ int gd_lane_id() {
  int l;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
  return l;
}
__device__ __forceinline__ float gd_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float gd_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float gd_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// hydrophobic polynomial 1 - (a r^2 - b r^4 + c r^6 - d r^8), r^2 in A^2
#define GD_HPA (3.5f / 81.0f)
#define GD_HPB (4.5f / 81.0f / 81.0f)
#define GD_HPC (2.5f / 81.0f / 81.0f / 81.0f)
#define GD_HPD (0.5f / 81.0f / 81.0f / 81.0f / 81.0f)
// log2(e) / 2 in FP64: the H-bond exponent is formed in double precision
#define GD_HALF_LOG2E_D 0.72134752044448170368

// ---------------------------------------------------------------------------
// Address-space helpers: tables shared by all lanes are read through the
// constant address space, which with a wave-uniform address compiles to
// s_load (scalar data cache, one fetch for 64 lanes); per-lane data goes
// through the global address space (global_load/store, not flat).
#if defined(__HIP_DEVICE_COMPILE__)
#define GD_AS_CONST __attribute__((address_space(4)))
#define GD_AS_GLOBAL __attribute__((address_space(1)))
#else  // host pass of this translation unit: kernels are only declared there
#define GD_AS_CONST
#define GD_AS_GLOBAL
#endif
template <class T>
__device__ __forceinline__ const GD_AS_CONST T *gd_c(const T *p) {
  return (const GD_AS_CONST T *)p;
}
template <class T>
__device__ __forceinline__ GD_AS_GLOBAL T *gd_g(T *p) {
  return (GD_AS_GLOBAL T *)p;
}
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ const GD_AS_GLOBAL T *gd_g(const T *p) {
  return (const GD_AS_GLOBAL T *)p;
}
#endif

// FP32 -> FP64 conversions that the compiler may not move (see the H-bond
// term); _s takes a wave-uniform value held in a scalar register.
__device__ __forceinline__ double gd_f2d(float x) {
  double d;
  asm volatile("v_cvt_f64_f32 %0, %1" : "=v"(d) : "v"(x));
  return d;
}
__device__ __forceinline__ double gd_f2d_s(float x) {
  double d;
  asm volatile("v_cvt_f64_f32 %0, %1" : "=v"(d) : "s"(x));
  return d;
}
// r2 += (a - b)^2 with the difference exact in FP64
__device__ __forceinline__ void gd_hdb_axis(float a, float b, double &r2) {
  const double d = gd_f2d(a) - gd_f2d(b);
  r2 = __builtin_fma(d, d, r2);
}

// ---------------------------------------------------------------------------
// Pair kernel.  LC ligand atoms per wave, all LC coordinates in registers.

// Per-item constants of the pair loop.
struct GdPairConst {
  float lj0, lj01;           // LJ denominator lj0 (1 + lj1 r^2) = lj0 + lj0 lj1 r^2
  float e0, e2, e3;          // el1 * (el0 + a1 s^2 + b1 s^3), s = el1 r, as a polynomial in r
  float inv_el1;
};

#define GD_ROW 9  // force-field row entries used by the pair loop
#define GD_ROW_BLOCK (GD_PTAB_PT * GD_PTAB_W * 4)  // bytes of one ligand type's row block

// One (atom, point) pair: the six terms of energy_prt_v1.C:30-134 for one
// ligand atom of the lane against one protein point, with the force-field row
// rw of (atom type, point type) in scalar registers.  CLS: bit 0 = cdc point
// (hydrophobic term), bit 1 = class-2 point (psp term); the caller dispatches
// on the point's flags once, so the pairs carry no per-pair class tests.
template <int CLS>
__device__ __forceinline__ void gd_pair_one(
    const GdPairConst &k, const float (&rw)[GD_ROW], const GD_AS_CONST float *__restrict__ pspv,
    const int P, const int p, const unsigned prow, const float4 c, const float pele, const float phpp,
    const float pmf_a, const float lxj, const float lyj, const float lzj, float &elej, float &hpcj,
    float &evdw, float &epmf, float &epsp, float &ehdb) {
    const float w9 = rw[0], nq = rw[1], pB = rw[2], pC = rw[3];
    const float pmf1 = rw[4], pthr = rw[5], hdb1 = rw[6], hdb0 = rw[7], r2far = rw[8];
    // the sign bit of the psp threshold flags an active H-bond pair; the asm
    // keeps the test on the scalar unit (as a float compare it became a VALU op)
    int hflag;
    asm("" : "=s"(hflag) : "0"(__float_as_int(pthr)));
    const float dx = lxj - c.x, dy = lyj - c.y, dz = lzj - c.z;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    const float ri = gd_rsq(r2);
    const float rr = r2 * ri;
    // soft Lennard-Jones (:61-64): (p1 - p2) / (p1 lj0 (1 + lj1 r^2) + 1) with
    // p1 = p1a / r^9, p2 = p2a / r^6, multiplied through by r^9 / p1a:
    // (1 - q r^3) / (lj0 (1 + lj1 r^2) + w r^9), q = p2a / p1a, w = 1 / p1a
    // (rows with p1a = p2a = 0 carry w = inf: the term is 0)
    {
      const float r3 = r2 * rr;
      const float r9 = (r3 * r3) * r3;
      const float num = __builtin_fmaf(nq, r3, 1.0f);
      const float den = __builtin_fmaf(w9, r9, __builtin_fmaf(k.lj01, r2, k.lj0));
      evdw = __builtin_fmaf(num, gd_rcp(den), evdw);
      // pinned here: sunk past the branches below, the terms of every atom of
      // the iteration stayed live to its end (registers)
      asm volatile("" : "+v"(evdw));
    }
    // Contact PMF (:88-91): pmf1 / (1 + exp2(x)), x = (6 - r/2)(r - pmf0) log2(e).
    // It is summed as pmf1 - pmf1 / (1 + exp2(-x)): the pmf1 of every pair of
    // the job is one constant added once per replica, and epmf collects only the
    // corrections.  Beyond the row's r^2 threshold x <= -26, so 1 + exp2(x)
    // rounds to 1: the reference term is exactly pmf1 and the correction is
    // taken as exactly 0, on every lane, so a replica's sum does not depend on
    // the other lanes of its wave.  Also s = el1 r >= 1 there, so the
    // electrostatic factor is 1/s (scaled: 1/r).  Far field, decided for the
    // whole wave on one compare: no correction, g = 1/r.
    float g = ri;
    const bool near = r2 < r2far;
    if (__ballot(near) != 0) {
      asm volatile("");
      // -x = log2(e)/2 r^2 + pB r + pC
      const float nx = __builtin_fmaf(pB, rr, __builtin_fmaf(pmf_a, r2, pC));
      const float t = gd_rcp(1.0f + gd_exp2(nx));
      epmf = __builtin_fmaf(-pmf1, near ? t : 0.0f, epmf);
      // distance-dependent electrostatics (:72-78), charge applied per atom:
      // el1 g = poly(r) for s = el1 r < 1, else 1/r.  s < 1 is a close
      // contact (r < 2 A with the published el1 = 0.5), so the polynomial is
      // evaluated only when some lane of the wave is that close
      const bool close = rr < k.inv_el1;
      if (__ballot(close) != 0) {
        asm volatile("");
        const float poly = __builtin_fmaf(__builtin_fmaf(k.e3, rr, k.e2), r2, k.e0);
        g = close ? poly : ri;
      }
      // The terms below vanish beyond the row's threshold too (it is raised
      // to cover their ranges), so far waves skip them.
      // hydrophobic contact (energy_prt_v1.C:44-50):
      // hpp (1 - r^2 (a - b r^2 + c r^4 - d r^6)) for r^2 <= 81, on the lanes
      // within 9 A only (exec-masked; skipped when no lane is).  The empty
      // volatile asm keeps the branch: without it the compiler evaluates the
      // term for every pair and selects
      if (CLS & 1) {
        if (r2 <= 81.0f) {
          asm volatile("");
          float w = __builtin_fmaf(-GD_HPD, r2, GD_HPC);
          w = __builtin_fmaf(w, r2, -GD_HPB);
          w = __builtin_fmaf(w, r2, GD_HPA);
          hpcj = __builtin_fmaf(phpp, __builtin_fmaf(-r2, w, 1.0f), hpcj);
        }
      }
      // pocket-specific potential (:100-107): r <= pmf0 decided on r^2
      // against the exact float threshold of sqrtf(r^2) <= pmf0; the value is
      // loaded only when some lane is in contact (exec-masked branch)
      if (CLS & 2) {
        if (r2 <= fabsf(pthr)) {
          asm volatile("");
          epsp += pspv[(size_t)prow * P + p];
        }
      }
      // hydrogen bond (:116-121), hdb1 exp(-h^2 / 2) with h = (r - hdb0) hdb1,
      // as hdb1 2^-q, q = log2(e)/2 h^2; pairs with hdb0 <= 0.1 skip the term
      // on a wave-uniform branch.  In the tail of the Gaussian the relative
      // error of a contribution is about q' dr, and FP32 rounding of r (and of
      // q itself, up to ~150) would put it near 1e-5; so q is formed in FP64:
      // exact coordinate differences, r from the FP32 r by one Newton step,
      // then 2^-q = 2^-floor(q) * 2^-frac(q) with only frac(q) in [0, 1)
      // rounded to FP32.  Beyond the row's r^2 threshold q >= 160, where the
      // FP32 result is exactly 0 (far waves skip it).
      if (hflag < 0) {
        asm volatile("");
        // the conversions are volatile asm so that they stay here: hoisted out
        // of the point loop, the converted ligand coordinates would hold six
        // VGPRs per atom; r^2 is accumulated one axis at a time
        double r2d = 0.0;
        gd_hdb_axis(lzj, c.z, r2d);
        gd_hdb_axis(lyj, c.y, r2d);
        gd_hdb_axis(lxj, c.x, r2d);
        const double r0 = gd_f2d(rr);
        // 1/(2 r) from rr, not from ri: ri is then dead once g is chosen, and
        // g shares its register (no copy on the far path)
        const double rd = __builtin_fma(__builtin_fma(-r0, r0, r2d), gd_f2d(0.5f * gd_rcp(rr)), r0);
        const double hd = (rd - gd_f2d_s(hdb0)) * gd_f2d_s(hdb1);
        const double q = (hd * hd) * GD_HALF_LOG2E_D;
        const float fq = (float)__builtin_amdgcn_fract(q);
        const int nq = (int)q;
        ehdb = __builtin_fmaf(hdb1, __builtin_ldexpf(gd_exp2(-fq), -nq), ehdb);
      }
    }
    elej = __builtin_fmaf(pele, g, elej);
}

// Rows of the LC chunk atoms for one point type, loaded into scalar registers.
template <int LC>
__device__ __forceinline__ void gd_load_rows(const GD_AS_CONST char *__restrict__ ptab,
                                             const unsigned (&trow)[LC], const unsigned poff,
                                             float (&rw)[LC][GD_ROW]) {
#pragma unroll
  for (int j = 0; j < LC; ++j) {
    const GD_AS_CONST float *__restrict__ prm =
        (const GD_AS_CONST float *)(ptab + (trow[j] * GD_ROW_BLOCK + poff));
#pragma unroll
    for (int i = 0; i < GD_ROW; ++i) rw[j][i] = prm[i];
  }
}

// The LC chunk atoms against one point, rows given (held across the points of
// one point type).
template <int LC, int CLS>
__device__ __forceinline__ void gd_pair_atoms_rows(
    const GdPairConst &k, const float (&rw)[LC][GD_ROW], const GD_AS_CONST float *__restrict__ pspv,
    const int P, const int p, const float4 c, const float pele, const float phpp, const float pmf_a,
    const float (&lx)[LC], const float (&ly)[LC], const float (&lz)[LC], const unsigned (&trow)[LC],
    float (&ele)[LC], float (&hpc)[LC], float &evdw, float &epmf, float &epsp, float &ehdb) {
#pragma unroll
  for (int j = 0; j < LC; ++j) {
    gd_pair_one<CLS>(k, rw[j], pspv, P, p, trow[j], c, pele, phpp, pmf_a, lx[j], ly[j], lz[j], ele[j],
                     hpc[j], evdw, epmf, epsp, ehdb);
    // one atom at a time: interleaving atoms raises register use for no gain
    // in issue rate
    __builtin_amdgcn_sched_barrier(0);
  }
}

// The LC chunk atoms against one point, each atom's row loaded for this point
// (wide chunks, whose rows do not fit the scalar registers).  The loads are
// software-pipelined: the row of atom j + 1 is loaded while atom j is
// evaluated.  Scalar loads complete out of order, so every wait is for all of
// them; the asm with a memory clobber places the wait for row j before the
// load of row j + 1 is issued.
template <int LC, int CLS>
__device__ __forceinline__ void gd_pair_atoms_load(
    const GdPairConst &k, const GD_AS_CONST char *__restrict__ ptab,
    const GD_AS_CONST float *__restrict__ pspv, const int P, const int p, const unsigned poff,
    const float4 c, const float pele, const float phpp, const float pmf_a, const float (&lx)[LC],
    const float (&ly)[LC], const float (&lz)[LC], const unsigned (&trow)[LC], float (&ele)[LC],
    float (&hpc)[LC], float &evdw, float &epmf, float &epsp, float &ehdb) {
  float rw[GD_ROW];
  {
    const GD_AS_CONST float *__restrict__ prm =
        (const GD_AS_CONST float *)(ptab + (trow[0] * GD_ROW_BLOCK + poff));
#pragma unroll
    for (int i = 0; i < GD_ROW; ++i) rw[i] = prm[i];
  }
#pragma unroll
  for (int j = 0; j < LC; ++j) {
    asm volatile("" ::"s"(rw[0]), "s"(rw[1]), "s"(rw[2]), "s"(rw[3]), "s"(rw[4]), "s"(rw[5]),
                 "s"(rw[6]), "s"(rw[7]), "s"(rw[8])
                 : "memory");
    float cur[GD_ROW];
#pragma unroll
    for (int i = 0; i < GD_ROW; ++i) cur[i] = rw[i];
    if (j + 1 < LC) {
      const GD_AS_CONST float *__restrict__ prm =
          (const GD_AS_CONST float *)(ptab + (trow[j + 1 < LC ? j + 1 : j] * GD_ROW_BLOCK + poff));
#pragma unroll
      for (int i = 0; i < GD_ROW; ++i) rw[i] = prm[i];
    }
    gd_pair_one<CLS>(k, cur, pspv, P, p, trow[j], c, pele, phpp, pmf_a, lx[j], ly[j], lz[j], ele[j],
                     hpc[j], evdw, epmf, epsp, ehdb);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// The points of one class [lo, hi) of one tile against the LC chunk atoms:
// tile b of nP takes lo + b, lo + b + nP, ...  The host lays each class out so
// that these are a contiguous piece of the class's points sorted by point
// type, so the force-field rows of a 4-atom chunk change only where the point
// type does and stay in scalar registers across the run; within a type the
// order spreads the distances to the pocket evenly over the tiles, so every
// tile holds the same mix of near and far points and waves cost the same.
template <int LC, int CLS>
__device__ __forceinline__ void gd_pair_class(
    const GD_AS_CONST GdJob &J, const int lo, const int hi, const int b, const unsigned pcoff,
    const GdPairConst &k, const float pmf_a,
    const float (&lx)[LC], const float (&ly)[LC], const float (&lz)[LC], const unsigned (&trow)[LC],
    float (&ele)[LC], float (&hpc)[LC], float &evdw, float &epmf, float &epsp, float &ehdb) {
  const int P = J.P, nP = J.nP;
  const int p0 = lo + b;
  if (p0 >= hi) return;
  const GD_AS_GLOBAL char *__restrict__ pcb = (const GD_AS_GLOBAL char *)gd_g(J.pcoord);
  const GD_AS_GLOBAL char *__restrict__ pab = (const GD_AS_GLOBAL char *)gd_g(J.pattr);
  const GD_AS_CONST char *__restrict__ ptab = (const GD_AS_CONST char *)gd_c(J.ptab);
  const GD_AS_CONST float *__restrict__ pspv = gd_c(J.pspv);
  // The point attributes (uniform) go through the vector memory path and are
  // prefetched one point ahead with the coordinates: as scalar loads they
  // stalled the wave.  Both loads take a uniform base in SGPRs plus a 32-bit
  // VGPR offset (zero for the attributes, the lane's protein conformer for the
  // coordinates), redeclared opaque at each load so that the compiler does not
  // fold base and offset into a 64-bit VGPR address: the point index then
  // costs scalar adds only, no per-point VALU.
  unsigned zo = 0, vo = pcoff;
  auto attr = [&](int q) {
    asm("" : "+v"(zo));
    return *(const GD_AS_GLOBAL float4 *)(pab + (size_t)q * 16 + zo);
  };
  auto coord = [&](int q) {
    asm("" : "+v"(vo));
    return *(const GD_AS_GLOBAL float4 *)(pcb + (size_t)q * 16 + vo);
  };
  constexpr bool kRows = LC == 4;  // rows of 4 atoms fit the scalar registers
  float rw[kRows ? LC : 1][GD_ROW];
  int ty_cur = -1;
  auto point = [&](int q, const float4 c, const float4 pa) {
    const int fl = __builtin_amdgcn_readfirstlane(__float_as_int(pa.z));
    const int ty = fl & GD_PF_TYPE;
    // byte offset of the point type's entry inside a ligand type's row block
    const unsigned poff = (unsigned)ty * (GD_PTAB_W * 4);
    if constexpr (kRows) {
      if (ty != ty_cur) {
        ty_cur = ty;
        gd_load_rows<LC>(ptab, trow, poff, *reinterpret_cast<float(*)[LC][GD_ROW]>(&rw));
      }
      gd_pair_atoms_rows<LC, CLS>(k, *reinterpret_cast<float(*)[LC][GD_ROW]>(&rw), pspv, P, q, c,
                                  pa.x, pa.y, pmf_a, lx, ly, lz, trow, ele, hpc, evdw, epmf, epsp,
                                  ehdb);
    } else {
      gd_pair_atoms_load<LC, CLS>(k, ptab, pspv, P, q, poff, c, pa.x, pa.y, pmf_a, lx, ly, lz, trow,
                                  ele, hpc, evdw, epmf, epsp, ehdb);
    }
  };
  // Two points per iteration, each prefetching the next point into the other
  // register set, so the loop carries no register copies.  The prefetch index
  // is clamped into the class (the last one reloads a point, unused).
  float4 c0 = coord(p0), a0 = attr(p0);
  int p = p0;
  for (; p + nP < hi; p += 2 * nP) {
    const float4 c1 = coord(p + nP), a1 = attr(p + nP);
    point(p, c0, a0);
    const int q = min(p + 2 * nP, hi - 1);
    c0 = coord(q);
    a0 = attr(q);
    point(p + nP, c1, a1);
  }
  if (p < hi) point(p, c0, a0);
}

// The point loop of one item, class by class (padding atoms of a partial
// chunk contribute exactly zero).
template <int LC>
__device__ __forceinline__ void gd_pair_points(
    const GD_AS_CONST GdJob &J, const int b, const unsigned pc,
    const GdPairConst &k, const float (&lx)[LC], const float (&ly)[LC], const float (&lz)[LC],
    const unsigned (&trow)[LC], float (&ele)[LC], float (&hpc)[LC], float &evdw, float &epmf,
    float &epsp, float &ehdb) {
  // log2(e)/2 held in a VGPR: a VALU op reads at most one SGPR or literal
  float pmf_a;
  asm("v_mov_b32 %0, 0x3f38aa3b" : "=v"(pmf_a));  // log2(e) / 2
  gd_pair_class<LC, 0>(J, J.pcls[0], J.pcls[1], b, pc, k, pmf_a, lx, ly, lz, trow, ele, hpc,
                       evdw, epmf, epsp, ehdb);
  gd_pair_class<LC, 1>(J, J.pcls[1], J.pcls[2], b, pc, k, pmf_a, lx, ly, lz, trow, ele, hpc,
                       evdw, epmf, epsp, ehdb);
  gd_pair_class<LC, 2>(J, J.pcls[2], J.pcls[3], b, pc, k, pmf_a, lx, ly, lz, trow, ele, hpc,
                       evdw, epmf, epsp, ehdb);
  gd_pair_class<LC, 3>(J, J.pcls[3], J.pcls[4], b, pc, k, pmf_a, lx, ly, lz, trow, ele, hpc,
                       evdw, epmf, epsp, ehdb);
}

template <int LC>
__device__ __forceinline__ void gd_pair_item(const GdJob *__restrict__ jobs, const GdItem it,
                                             const int lane, const int wave,
                                             float (*__restrict__ red)[GD_NLIN + LC][64]) {
  const GD_AS_CONST GdJob &J = gd_c(jobs)[it.job];
  const int Rp = J.Rp, P = J.P, L = J.L, nP = J.nP;
  const int b = it.b * GD_GROUP + wave;  // this wave's protein tile
  const int r = it.rb + lane;  // < Rp by construction (Rp padded to 64)
  const int l0 = it.a * LC;
  const int nl = min(LC, L - l0);
  const GD_AS_CONST GdAtom *__restrict__ atoms = gd_c(J.atoms);
  const GD_AS_GLOBAL float *__restrict__ coords = gd_g(J.coords);

  // ligand chunk: coordinates of this lane's replica, per-atom constants (uniform)
  float lx[LC], ly[LC], lz[LC], ele[LC], hpc[LC];
  unsigned trow[LC];  // row block of each atom
#pragma unroll
  for (int j = 0; j < LC; ++j) {
    const int l = l0 + min(j, nl - 1);
    lx[j] = coords[(size_t)(3 * l + 0) * Rp + r];
    ly[j] = coords[(size_t)(3 * l + 1) * Rp + r];
    lz[j] = coords[(size_t)(3 * l + 2) * Rp + r];
    trow[j] = (unsigned)atoms[l0 + j].tidx;
    ele[j] = 0.0f;
    hpc[j] = 0.0f;
  }
  // byte offset of this lane's protein conformer in the point coordinates
  const unsigned pc = (unsigned)gd_g(J.prti)[r] * (unsigned)P * 16u;
  GdPairConst k;
  k.lj0 = J.lj0;
  k.lj01 = J.lj0 * J.lj1;
  const float el1 = J.el1;
  k.e0 = J.el0 * el1;
  k.e2 = J.a1 * el1 * el1 * el1;
  k.e3 = J.b1 * el1 * el1 * el1 * el1;
  k.inv_el1 = J.inv_el1;
  const float inv_el1 = k.inv_el1;
  float evdw = 0.0f, epmf = 0.0f, epsp = 0.0f, ehdb = 0.0f;
  gd_pair_points<LC>(J, b, pc, k, lx, ly, lz, trow, ele, hpc, evdw, epmf, epsp, ehdb);
  // the lane's replica for the stores below, derived again rather than kept
  // live through the point loop (where it would take two VGPRs)
  const int lane_e = gd_lane_id();
  const int r_e = it.rb + lane_e;

  // KDE: this tile's share of every chunk atom's anchor bin, a contiguous range
  // [f(b), f(b+1)) with f(b) = floor(n b / nP) evaluated in FP32: f is
  // monotonic, f(0) = 0 and the last tile ends at n, so the tiles partition the
  // bin whatever the rounding (contiguous ranges keep the scalar loads in cache)
  float ekde = 0.0f;
  const float kde2l = J.kde2l;
  const GD_AS_CONST float4 *__restrict__ kde = gd_c(J.kde);
  const float inv_np = 1.0f / (float)nP;
#pragma unroll
  for (int j = 0; j < LC; ++j) {
    {
      const GdAtom at = atoms[l0 + j];  // padding atoms: no bins
      const int nb = at.kde_e - at.kde_b;
      const int k0 = at.kde_b + gd_uniform((int)((float)(nb * b) * inv_np));
      const int k1 = b + 1 == nP ? at.kde_e : at.kde_b + gd_uniform((int)((float)(nb * (b + 1)) * inv_np));
      float s = 0.0f;
      for (int k = k0; k < k1; ++k) {
        const float4 kp = kde[k];
        const float dx = lx[j] - kp.x, dy = ly[j] - kp.y, dz = lz[j] - kp.z;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        s += gd_exp2(kde2l * d2);
      }
      ekde = __builtin_fmaf(s, at.kde_inv_n, ekde);
    }
    __builtin_amdgcn_sched_barrier(0);
  }

  // MCS: this tile's rows b, b + nP, ..., restricted to the chunk atoms
  {
    const int M = J.M, nL = J.nL;
    const GD_AS_CONST float4 *__restrict__ mcs = gd_c(J.mcs);
    GD_AS_GLOBAL float *__restrict__ pm = gd_g(J.part_mcs);
    for (int m = b; m < M; m += nP) {
      float d = 0.0f;
#pragma unroll
      for (int j = 0; j < LC; ++j) {
        const float4 t = mcs[(size_t)m * J.Ls + l0 + j];  // padding atoms: w == 0
        if (t.w != 0.0f) {
          const float dx = lx[j] - t.x, dy = ly[j] - t.y, dz = lz[j] - t.z;
          d += __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        }
      }
      pm[((size_t)m * nL + it.a) * Rp + r_e] = d;
    }
  }

  // partial sums
  float eele = 0.0f;
#pragma unroll
  for (int j = 0; j < LC; ++j) eele = __builtin_fmaf(atoms[l0 + j].q, ele[j], eele);  // padding: q = 0
  eele *= inv_el1;
  // group reduction: the group's tiles summed in wave order
  red[wave][0][lane_e] = evdw;
  red[wave][1][lane_e] = eele;
  red[wave][2][lane_e] = epmf;
  red[wave][3][lane_e] = epsp;
  red[wave][4][lane_e] = ehdb;
  red[wave][5][lane_e] = ekde;
#pragma unroll
  for (int j = 0; j < LC; ++j) red[wave][GD_NLIN + j][lane_e] = hpc[j];
  __syncthreads();
  const int nG = J.nG;
  const size_t nS = (size_t)J.nL * nG;
  const int slot = it.a * nG + it.b;
  GD_AS_GLOBAL float *__restrict__ pl = gd_g(J.part_lin);
  GD_AS_GLOBAL float *__restrict__ ph = gd_g(J.part_hpc);
  for (int q = wave; q < GD_NLIN + LC; q += GD_GROUP) {
    float v = red[0][q][lane_e];
#pragma unroll
    for (int w = 1; w < GD_GROUP; ++w) v += red[w][q][lane_e];
    if (q < GD_NLIN)
      pl[((size_t)q * nS + slot) * Rp + r_e] = v;
    else if (q - GD_NLIN < nl)
      ph[((size_t)(l0 + q - GD_NLIN) * nG + it.b) * Rp + r_e] = v;
  }
  __syncthreads();  // red[] is reused by the next item of a persistent group
}

// One work group of GD_GROUP waves per item (grid = items).
template <int LC>
__global__ __launch_bounds__(64 * GD_GROUP) __attribute__((amdgpu_waves_per_eu(LC == 4 ? 8 : 1))) void gd_pair_kernel(const GdJob *__restrict__ jobs,
                                                       const GdItem *__restrict__ items,
                                                       int n_items) {
  __shared__ float red[GD_GROUP][GD_NLIN + LC][64];
  const int iid = gd_uniform((int)blockIdx.x);
  if (iid >= n_items) return;  // whole group: no wave waits at the barriers
  gd_pair_item<LC>(jobs, gd_c(items)[iid], threadIdx.x & 63, gd_uniform((int)(threadIdx.x >> 6)), red);
}

// Persistent form: a fixed grid sized to the machine; every group pulls item
// indices from a device counter until the list is exhausted, which balances
// items of uneven cost (complexes of different sizes in one launch).  The
// counter is zeroed by the kernel that runs before each launch (finish or
// transform kernel); every group exits once the counter passes n_items.
template <int LC>
__global__ __launch_bounds__(64 * GD_GROUP) __attribute__((amdgpu_waves_per_eu(LC == 4 ? 8 : 1))) void gd_pair_kernel_wq(const GdJob *__restrict__ jobs,
                                                          const GdItem *__restrict__ items,
                                                          int n_items, int *__restrict__ wq) {
  __shared__ float red[GD_GROUP][GD_NLIN + LC][64];
  __shared__ int s_iid;
  const int lane = threadIdx.x & 63, wave = gd_uniform((int)(threadIdx.x >> 6));
  for (;;) {
    if (threadIdx.x == 0) s_iid = atomicAdd(wq, 1);
    __syncthreads();
    const int iid = gd_uniform(s_iid);
    if (iid >= n_items) break;  // the same value in every wave: the group leaves together
    // s_iid is rewritten only after the item's barriers, which every wave
    // reaches after reading it
    gd_pair_item<LC>(jobs, gd_c(items)[iid], lane, wave, red);
  }
}

template __global__ void gd_pair_kernel<4>(const GdJob *, const GdItem *, int);
template __global__ void gd_pair_kernel<8>(const GdJob *, const GdItem *, int);
template __global__ void gd_pair_kernel<16>(const GdJob *, const GdItem *, int);
template __global__ void gd_pair_kernel_wq<4>(const GdJob *, const GdItem *, int, int *);
template __global__ void gd_pair_kernel_wq<8>(const GdJob *, const GdItem *, int, int *);
template __global__ void gd_pair_kernel_wq<16>(const GdJob *, const GdItem *, int, int *);

// ---------------------------------------------------------------------------
// Ligand transform of one replica's proposal (kernel_cpp_montecarlo_soa.C:
// 256-293): waves of the block split the atoms, lanes are replicas.
// Every product of the rotation is rounded on its own (no contraction), and
// the transform below is one explicit FMA chain, so that every kernel that
// transforms a pose (transform, finish, swap) gives the same bits for it.
__device__ __forceinline__ void gd_rotation(const float m[6], float R[9]) {
#pragma clang fp contract(off)
  float s1, c1, s2, c2, s3, c3;
  sincosf(m[3], &s1, &c1);
  sincosf(m[4], &s2, &c2);
  sincosf(m[5], &s3, &c3);
  R[0] = c1 * c2;
  R[1] = c1 * s2 * s3 - c3 * s1;
  R[2] = s1 * s3 + c1 * c3 * s2;
  R[3] = c2 * s1;
  R[4] = c1 * c3 + s1 * s2 * s3;
  R[5] = c3 * s1 * s2 - c1 * s3;
  R[6] = -s2;
  R[7] = c2 * s3;
  R[8] = c2 * c3;
}

// The finish, transform and move code reads the job struct through the
// constant address space (uniform fields become scalar loads) and per-replica
// arrays through the global one; generic pointers made every access a flat
// load followed by a full wait.
typedef const GD_AS_CONST GdJob GdJobC;

// One atom of the transform x = R a + t (kernel_cpp_montecarlo_soa.C:274-293),
// the conformer-relative products summed first and the lab-frame offset t
// added last, into the [L][3][Rp] coordinate layout.
__device__ __forceinline__ void gd_put_atom(GD_AS_GLOBAL float *co, int Rp, int r, int l,
                                            const float R[9], const float t[3], const float4 a) {
  co[(size_t)(3 * l + 0) * Rp + r] = __builtin_fmaf(R[0], a.x, __builtin_fmaf(R[1], a.y, R[2] * a.z)) + t[0];
  co[(size_t)(3 * l + 1) * Rp + r] = __builtin_fmaf(R[3], a.x, __builtin_fmaf(R[4], a.y, R[5] * a.z)) + t[1];
  co[(size_t)(3 * l + 2) * Rp + r] = __builtin_fmaf(R[6], a.x, __builtin_fmaf(R[7], a.y, R[8] * a.z)) + t[2];
}

__device__ __forceinline__ void gd_transform_atoms(GdJobC &J, int r, const float R[9],
                                                   const float t[3], int wave, int nwave) {
  const int L = J.L, Rp = J.Rp;
  const GD_AS_GLOBAL float4 *lc = gd_g(J.lig0) + (size_t)gd_g(J.ligi)[r] * L;
  GD_AS_GLOBAL float *co = gd_g(J.coords);
  for (int l = wave; l < L; l += nwave) gd_put_atom(co, Rp, r, l, R, t, lc[l]);
}

// Transform the current proposal (prop) of every replica in the block.
__global__ __launch_bounds__(256) void gd_transform_kernel(const GdJob *__restrict__ jobs,
                                                            const GdBlock *__restrict__ blocks,
                                                            int *__restrict__ wq) {
  if (wq && blockIdx.x == 0 && threadIdx.x == 0) *wq = 0;  // next pair launch's queue
  const GdBlock bk = gd_c(blocks)[blockIdx.x];
  GdJobC &J = gd_c(jobs)[bk.job];
  const int lane = threadIdx.x & 63, wave = gd_uniform((int)(threadIdx.x >> 6));
  const int r = bk.rb + lane;
  const GD_AS_GLOBAL float *prop = gd_g(J.prop);
  float m[6];
  for (int i = 0; i < 6; ++i) m[i] = prop[(size_t)i * J.Rp + r];
  float R[9];
  gd_rotation(m, R);
  const float t[3] = {m[0] + J.center[0], m[1] + J.center[1], m[2] + J.center[2]};
  gd_transform_atoms(J, r, R, t, wave, 4);
}

// ---------------------------------------------------------------------------
// Move increments for step `step` (reference: move_scale * U[0,1),
// kernel_cpp_montecarlo_soa.C:242-250; gaussian mode: move_scale * N(0,1)).
__device__ __forceinline__ void gd_move(GdJobC &J, int r, uint32_t step, float inc[6],
                                        float *u_acc) {
  uint32_t w[8];
  gd_step_draws(((uint64_t)J.seed_hi << 32) | J.seed_lo, J.job_id, (uint32_t)r, step, w);
  float scale[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) scale[i] = J.move_scale[i];
  gd_move_incr(J.move_mode, scale, w, inc);
  *u_acc = gd_u01(w[6]);
}

// Increments of replica-steps (replica r of n_rep, step step0 + s), one per
// thread: the draws gd_move makes, exposed for the statistical tests.
__global__ __launch_bounds__(256) void gd_move_draws_kernel(uint64_t seed, uint32_t job, int mode,
                                                            GdScale6 scale, int n_rep,
                                                            uint32_t step0, int n,
                                                            float *__restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t w[8];
  gd_step_draws(seed, job, (uint32_t)(i % n_rep), step0 + (uint32_t)(i / n_rep), w);
  float inc[6];
  gd_move_incr(mode, scale.v, w, inc);
#pragma unroll
  for (int k = 0; k < 6; ++k) out[(size_t)i * 6 + k] = inc[k];
}

// ---------------------------------------------------------------------------
// Sum of n values at stride `st` (floats) with eight loads in flight, the last
// partial block too (indices clamped to n - 1, masked to zero); element q goes
// to accumulator q mod 8, so the order depends only on n and is reproducible.
__device__ __forceinline__ float gd_sum_strided(const GD_AS_GLOBAL float *__restrict__ p, int n,
                                                size_t st) {
  float a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = 0.0f;
  for (int q = 0; q < n; q += 8) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = p[(size_t)min(q + i, n - 1) * st];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] += q + i < n ? v[i] : 0.0f;
  }
  return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// Common shapes of the reduce kernel: at most QL linear slots, one atom and one
// MCS row per wave, at most QN tile groups and chunks.  Every load of the wave
// is issued before the first sum, one wait in all.  Indices are clamped into
// range and the extra values masked to zero; the sums run in slot order, so
// the result does not depend on QL or QN.
template <int QL, int QN>
__device__ __forceinline__ void gd_reduce_common(
    const GD_AS_GLOBAL float *__restrict__ part_lin, const GD_AS_GLOBAL float *__restrict__ part_hpc,
    const GD_AS_GLOBAL float *__restrict__ part_mcs, const GD_AS_CONST GdAtom *__restrict__ atoms,
    const GD_AS_CONST float2 *__restrict__ mcsrow, const int nS, const int nG, const int nL, const int L,
    const int M, const int Rp, const int r, const int wi, const int q0, const int q1, float (&acc)[8]) {
  const int ql = max(q1 - 1, q0), l = min(wi, L - 1), m = min(wi, max(M - 1, 0));
  float vl[GD_NLIN][QL], vh[QN], vm[QN];
#pragma unroll
  for (int c = 0; c < GD_NLIN; ++c)
#pragma unroll
    for (int i = 0; i < QL; ++i) vl[c][i] = part_lin[((size_t)c * nS + min(q0 + i, ql)) * Rp + r];
#pragma unroll
  for (int i = 0; i < QN; ++i) vh[i] = part_hpc[((size_t)l * nG + min(i, nG - 1)) * Rp + r];
#pragma unroll
  for (int i = 0; i < QN; ++i) vm[i] = part_mcs[((size_t)m * nL + min(i, nL - 1)) * Rp + r];
#pragma unroll
  for (int c = 0; c < GD_NLIN; ++c)
#pragma unroll
    for (int i = 0; i < QL; ++i) acc[c] += q0 + i < q1 ? vl[c][i] : 0.0f;
  float h = 0.0f, d = 0.0f;
#pragma unroll
  for (int i = 0; i < QN; ++i) {
    h += i < nG ? vh[i] : 0.0f;
    d += i < nL ? vm[i] : 0.0f;
  }
  if (wi < L) {  // hpc per atom (energy_prt_v1.C:132-133)
    const GdAtom at = atoms[l];
    const float t = (h - at.hpl0) / at.hpl1;
    acc[6] += 0.5f * t * t - at.hpl2;
  }
  if (wi < M) {  // MCS rows (energy_mcs_v1.C:5-29)
    const float2 row = mcsrow[m];
    acc[7] += row.x * sqrtf(d / row.y);
  }
}

// ---------------------------------------------------------------------------
// Reduce kernel: GD_RK work groups (4 waves) per 64-replica block; lanes are
// replicas.  The 4 GD_RK waves of a block split the partial slots: each takes a
// contiguous share of every linear term's slots, atoms l = w, w + 4 GD_RK, ...
// (the hpc sum over tile groups, then the non-linear restraint) and MCS rows
// m = w, w + 4 GD_RK, ... (the sum over chunks, then tcc sqrt(d/n)).  A group
// sums its 4 waves in a fixed order and writes 8 values per replica; the
// finish kernel sums the GD_RK slices in a fixed order.  The reduction reads
// most of the step's partials, so it is spread over many groups: one group per
// block kept too few loads in flight.
__global__ __launch_bounds__(256) void gd_reduce_kernel(const GdJob *__restrict__ jobs,
                                                         const GdBlock *__restrict__ blocks) {
  __shared__ float red[4][8][64];
  const int bi = blockIdx.x / GD_RK, k = blockIdx.x % GD_RK;
  const GdBlock bk = gd_c(blocks)[bi];
  GdJobC &J = gd_c(jobs)[bk.job];
  const int lane = threadIdx.x & 63, wave = gd_uniform((int)(threadIdx.x >> 6));
  const int r = bk.rb + lane;
  const int Rp = J.Rp;
  const int W = 4 * GD_RK, wi = 4 * k + wave;
  const int nS = J.nL * J.nG;
  const GD_AS_GLOBAL float *part_lin = gd_g(J.part_lin);
  const GD_AS_GLOBAL float *part_hpc = gd_g(J.part_hpc);
  const GD_AS_GLOBAL float *part_mcs = gd_g(J.part_mcs);
  const GD_AS_CONST GdAtom *atoms = gd_c(J.atoms);
  const GD_AS_CONST float2 *mcsrow = gd_c(J.mcsrow);
  float acc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) acc[c] = 0.0f;
  const int q0 = (int)(((long long)nS * wi) / W), q1 = (int)(((long long)nS * (wi + 1)) / W);
  const int nG = J.nG, nL = J.nL, L = J.L, M = J.M;
  if ((nS + W - 1) / W <= 2 && nG <= 12 && nL <= 12 && L <= W && M <= W) {
    gd_reduce_common<2, 12>(part_lin, part_hpc, part_mcs, atoms, mcsrow, nS, nG, nL, L, M, Rp, r, wi, q0,
                            q1, acc);
  } else if (q1 - q0 <= 4 && nG <= 16 && nL <= 16 && L <= W && M <= W) {
    gd_reduce_common<4, 16>(part_lin, part_hpc, part_mcs, atoms, mcsrow, nS, nG, nL, L, M, Rp, r, wi, q0,
                            q1, acc);
  } else {
#pragma unroll
    for (int c = 0; c < GD_NLIN; ++c)
      acc[c] = gd_sum_strided(part_lin + ((size_t)c * nS + q0) * Rp + r, q1 - q0, (size_t)Rp);
    for (int l = wi; l < L; l += W) {  // hpc per atom (energy_prt_v1.C:132-133)
      const float h = gd_sum_strided(part_hpc + (size_t)l * nG * Rp + r, nG, (size_t)Rp);
      const GdAtom at = atoms[l];
      const float t = (h - at.hpl0) / at.hpl1;
      acc[6] += 0.5f * t * t - at.hpl2;
    }
    for (int m = wi; m < M; m += W) {  // MCS rows (energy_mcs_v1.C:5-29)
      const float d = gd_sum_strided(part_mcs + (size_t)m * nL * Rp + r, nL, (size_t)Rp);
      const float2 row = mcsrow[m];
      acc[7] += row.x * sqrtf(d / row.y);
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) red[wave][c][lane] = acc[c];
  __syncthreads();
  if (wave == 0) {
    GD_AS_GLOBAL float *fp = gd_g(J.fin_part) + ((size_t)(bk.rb / 64) * GD_RK + k) * 8 * 64;
#pragma unroll
    for (int c = 0; c < 8; ++c)
      fp[c * 64 + lane] = ((red[0][c][lane] + red[1][c][lane]) + red[2][c][lane]) + red[3][c][lane];
  }
}

// ---------------------------------------------------------------------------
// Finish kernel: 1024 threads = 16 waves per 64-replica block; lanes are
// replicas.  Sums the reduce kernel's slices, assembles and weights the terms,
// records, accepts, proposes the next move and transforms the atoms.  Only the
// acceptance waits for the energies: the next step's increments, this step's
// accept draw and the rotations of both possible next proposals (from the
// proposal if it is accepted, from the current pose if not) are computed by
// waves 1-3 while the slices are summed, and wave 0 only selects.
#define GD_FW 16
#define GD_FPA 4  // conformer atoms per wave prefetched for the transform
__global__ __launch_bounds__(1024) void gd_finish_kernel(const GdJob *__restrict__ jobs,
                                                          const GdBlock *__restrict__ blocks,
                                                          int mode, int launch_step,
                                                          const int *__restrict__ step_base,
                                                          int *__restrict__ wq) {
  if (wq && blockIdx.x == 0 && threadIdx.x == 0) *wq = 0;  // next pair launch's queue
  if (step_base) launch_step += *step_base;  // graph replays: base step in device memory
  __shared__ float red[8][64];
  __shared__ float xf[2][12][64];  // next transform if this step is accepted [0] / rejected [1]
  __shared__ float inc_s[6][64];
  __shared__ float u_s[64];
  __shared__ int acc_s[64];
  const GdBlock bk = gd_c(blocks)[blockIdx.x];
  GdJobC &J = gd_c(jobs)[bk.job];
  const int lane = threadIdx.x & 63, wave = gd_uniform((int)(threadIdx.x >> 6));
  const int r = bk.rb + lane;
  const int Rp = J.Rp;
  const int step = launch_step - J.step0;  // this job's MC step
  const bool move = mode != GD_MODE_EVAL;

  // (1) the reduce kernel's slices: wave c < 8 sums value c over the GD_RK
  // slices in slice order (all loads issued first)
  if (wave < 8) {
    const GD_AS_GLOBAL float *fp =
        gd_g(J.fin_part) + (size_t)(bk.rb / 64) * GD_RK * 8 * 64 + (size_t)wave * 64 + lane;
    float v[GD_RK];
#pragma unroll
    for (int k = 0; k < GD_RK; ++k) v[k] = fp[(size_t)k * 8 * 64];
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < GD_RK; ++k) acc += v[k];
    red[wave][lane] = acc;
  }
  // (2) off the acceptance path: wave 1 the next step's increments, wave 2 this
  // step's accept draw (the initial pass draws u = 0, a forced accept)
  const uint32_t next = mode == GD_MODE_STEP ? (uint32_t)step + 1u : 0u;
  if (move && wave == 1) {
    float inc[6], un;
    gd_move(J, r, next, inc, &un);
#pragma unroll
    for (int i = 0; i < 6; ++i) inc_s[i][lane] = inc[i];
  }
  if (move && wave == 2)
    u_s[lane] = mode == GD_MODE_STEP
                    ? gd_accept_u01(((uint64_t)J.seed_hi << 32) | J.seed_lo, J.job_id, (uint32_t)r,
                                    (uint32_t)step)
                    : 0.0f;
  // the two possible bases of the next proposal, read before wave 0 writes them
  float base[6];
  if (move && (wave == 1 || wave == 3)) {
    const GD_AS_GLOBAL float *src = gd_g(wave == 1 ? J.prop : J.mm);
#pragma unroll
    for (int i = 0; i < 6; ++i) base[i] = src[(size_t)i * Rp + r];
  }
  // loads that do not wait for the energies, issued now: wave 0's replica
  // state, and every wave's first GD_FPA conformer atoms for the transform
  const GD_AS_GLOBAL float *prop_c = gd_g(J.prop), *en_c = gd_g(J.en), *mm_c = gd_g(J.mm);
  float m[6], m0[6], en9 = 0.0f, mbv = 0.0f, best0 = 0.0f;
  int nacc0 = 0;
  if (wave == 0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) m[i] = prop_c[(size_t)i * Rp + r];
    if (move) {
      en9 = en_c[(size_t)9 * Rp + r];
      mbv = gd_g(J.minus_beta)[gd_g(J.tmpi)[r]];
#pragma unroll
      for (int i = 0; i < 6; ++i) m0[i] = mm_c[(size_t)i * Rp + r];
      best0 = gd_g(J.best)[r];
      nacc0 = gd_g(J.nacc)[r];
    }
  }
  const int L = J.L;
  const GD_AS_GLOBAL float4 *lc = gd_g(J.lig0) + (size_t)gd_g(J.ligi)[r] * L;
  float4 pa[GD_FPA];
  if (move)
#pragma unroll
    for (int q = 0; q < GD_FPA; ++q) pa[q] = lc[min(wave + q * GD_FW, L - 1)];
  __syncthreads();

  if (move && (wave == 1 || wave == 3)) {  // rotation of each candidate next proposal
    float m[6], R[9];
#pragma unroll
    for (int i = 0; i < 6; ++i) m[i] = inc_s[i][lane] + base[i];
    gd_rotation(m, R);
    const int c = wave == 1 ? 0 : 1;
#pragma unroll
    for (int i = 0; i < 9; ++i) xf[c][i][lane] = R[i];
    xf[c][9][lane] = m[0] + J.center[0];
    xf[c][10][lane] = m[1] + J.center[1];
    xf[c][11][lane] = m[2] + J.center[2];
  }
  if (wave == 0) {
    GD_AS_GLOBAL float *prop = gd_g(J.prop), *en = gd_g(J.en), *mm = gd_g(J.mm);
    float s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = red[k][lane];
    // e_s assembly (kernel_cpp_montecarlo_soa.C:318-324), KDE/MCS/DST (:328-359)
    float e[10];
    e[0] = s[0];
    e[1] = s[1];
    e[2] = J.pmf_all + s[2];
    e[3] = s[3];
    e[4] = s[4] * (-1.0f / sqrtf(2.0f * 3.1415926535f));
    e[5] = s[6];
    e[6] = s[5] * J.kde3_inv;
    e[7] = logf(s[7] / (float)J.M);
    {
      float d2 = 0.0f;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        // (center + m) - pocket restated as m + (center - pocket): the loaders
        // put the ligand centre on the pocket centre, so this is m exactly, where
        // the reference's order rounds m to the ulp of the centre coordinate
        const float d = m[i] + J.cp[i];
        d2 += d * d;
      }
      e[8] = sqrtf(d2);
    }
    float raw[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) raw[i] = e[i];
    // normalisation + weights (:363-374)
#pragma unroll
    for (int i = 0; i < 7; ++i) e[i] = e[i] * J.inv_L;
    float tot = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      e[i] = J.a_para[i] * e[i] + J.b_para[i];
      tot += J.w[i] * e[i];
    }
    e[9] = tot;

    const bool live = r < J.R;
    if (!move) {
      if (live) {
        GD_AS_GLOBAL float *evt = gd_g(J.ev_terms), *evr = gd_g(J.ev_raw);
        for (int i = 0; i < 10; ++i) evt[(size_t)r * 10 + i] = e[i];
        for (int i = 0; i < 9; ++i) evr[(size_t)r * 9 + i] = raw[i];
      }
    } else {
      GD_AS_GLOBAL int *acc_r = gd_g(J.acc);
      // record the pre-move state if the previous step was accepted (:224-231)
      // into the ring half of this step's dump (dump d = step / rec_cap fills
      // half d mod 2 while the host copies out the other)
      if (mode == GD_MODE_STEP && J.record && live && acc_r[r]) {
        const int half = (step / J.rec_cap) & 1;
        GD_AS_GLOBAL int *rec_n = gd_g(J.rec_n) + (size_t)half * Rp;
        const int n = rec_n[r];
        if (n < J.rec_cap) {
          GD_AS_GLOBAL float *o =
              gd_g(J.rec) + (((size_t)half * Rp + r) * J.rec_cap + n) * GD_REC_W;
          o[0] = (float)step;
          for (int i = 0; i < 6; ++i) o[1 + i] = mm[(size_t)i * Rp + r];
          for (int i = 0; i < 10; ++i) o[7 + i] = en[(size_t)i * Rp + r];
          rec_n[r] = n + 1;
        }
      }
      // Metropolis (:382-392); the initial pass draws u = 0
      const float u = u_s[lane];
      const float x = (tot - en9) * mbv;
      const bool a = (u == 0.0f) ? (x >= GD_EXP_UNDERFLOW) : (u < expf(x));
      acc_s[lane] = live && a;  // padding lanes keep their pose
      if (live) {
        acc_r[r] = a;
        if (a) {
          for (int i = 0; i < 10; ++i) en[(size_t)i * Rp + r] = e[i];
          for (int i = 0; i < 6; ++i) mm[(size_t)i * Rp + r] = m[i];
          if (mode == GD_MODE_STEP) gd_g(J.nacc)[r] = nacc0 + 1;
          gd_g(J.best)[r] = fminf(best0, tot);
        }
      }
      // the next step's proposal, from the accepted pose (padding lanes keep
      // theirs: they never accept)
      const bool took = live && a;
#pragma unroll
      for (int i = 0; i < 6; ++i) prop[(size_t)i * Rp + r] = inc_s[i][lane] + (took ? m[i] : m0[i]);
    }
  }
  if (!move) return;
  __syncthreads();
  const int c = acc_s[lane] ? 0 : 1;
  float R[9], t[3];
#pragma unroll
  for (int i = 0; i < 9; ++i) R[i] = xf[c][i][lane];
  t[0] = xf[c][9][lane];
  t[1] = xf[c][10][lane];
  t[2] = xf[c][11][lane];
  // the transform of gd_transform_atoms, the first atoms from the prefetch
  GD_AS_GLOBAL float *co = gd_g(J.coords);
  auto put = [&](int l, const float4 a) { gd_put_atom(co, Rp, r, l, R, t, a); };
#pragma unroll
  for (int q = 0; q < GD_FPA; ++q)
    if (wave + q * GD_FW < L) put(wave + q * GD_FW, pa[q]);
  for (int l = wave + GD_FPA * GD_FW; l < L; l += GD_FW) put(l, lc[l]);
}

// ---------------------------------------------------------------------------
// Temperature exchange: one workgroup per (job, round) entry, threads over
// the (protein conformer, ligand conformer) ladders; pairs (t, t+1) for
// t = round mod 2, t+2, ...  Swaps temperature labels, not coordinates.
__global__ __launch_bounds__(256) void gd_exchange_kernel(const GdJob *__restrict__ jobs,
                                                           const GdExch *__restrict__ ex) {
  const GdExch e = ex[blockIdx.x];
  const GdJob &J = jobs[e.job];
  const int nt = J.n_tmp, Rp = J.Rp;
  const uint64_t seed = ((uint64_t)J.seed_hi << 32) | J.seed_lo;
  for (int ladder = threadIdx.x; ladder < J.n_prt * J.n_lig; ladder += blockDim.x) {
    int *own = J.own + (size_t)ladder * nt;
    for (int t = e.round & 1; t + 1 < nt; t += 2) {
      const int i = own[t], j = own[t + 1];
      const float u =
          gd_exchange_u01(seed, J.job_id, (uint32_t)(ladder * nt + t), (uint32_t)e.round);
      const float x = (J.minus_beta[t + 1] - J.minus_beta[t]) *
                      (J.en[(size_t)9 * Rp + i] - J.en[(size_t)9 * Rp + j]);
      if (u < expf(x)) {
        own[t] = j;
        own[t + 1] = i;
        J.tmpi[i] = t + 1;
        J.tmpi[j] = t;
        J.nexc[i] += 1;
        J.nexc[j] += 1;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Conformer pose swaps (Hamiltonian replica exchange along the ligand- or
// protein-conformer axis).  In round k, along axis `axis`, the lines of
// replicas that share the other conformer index and the temperature pair
// conformers (c, c+1) for c = k mod 2, c+2, ...; the two replicas of a pair
// propose each other's pose.  One swap sub-step is
//   gd_swap_propose_kernel  (prop := partner's pose, transform)
//   pair + reduce + finish(EVAL)  (energy of every replica at that pose)
//   gd_swap_accept_kernel   (joint Metropolis decision, next step's proposal)
// A job takes part when swr[job] >= 0 (the round) and its exchange axes hold
// `axis`; otherwise both kernels leave it untouched, and the evaluation in
// between changes nothing of its state.

// Partner of replica r along `axis` in round `rnd`, or -1; *lower is set when
// r holds the lower conformer index of the pair.
__device__ __forceinline__ int gd_swap_partner(GdJobC &J, int r, int axis, int rnd, bool *lower) {
  *lower = false;
  if (r >= J.R) return -1;
  const int l = gd_g(J.ligi)[r], p = gd_g(J.prti)[r], t = gd_g(J.tmpi)[r];
  const int n = axis ? J.n_prt : J.n_lig, c = axis ? p : l;
  const int c0 = c - (rnd & 1);
  if (c0 < 0) return -1;
  const int cq = (c0 & 1) ? c - 1 : c + 1;
  if (cq >= n) return -1;
  *lower = c < cq;
  const int ladder = axis ? cq * J.n_lig + l : p * J.n_lig + cq;
  return gd_g(J.own)[(size_t)ladder * J.n_tmp + t];
}

__device__ __forceinline__ bool gd_swap_active(GdJobC &J, int rnd, int axis) {
  return rnd >= 0 && ((J.xaxes >> (1 + axis)) & 1);  // GD_EXCH_LIG = 2, GD_EXCH_PRT = 4
}

__global__ __launch_bounds__(256) void gd_swap_propose_kernel(const GdJob *__restrict__ jobs,
                                                               const GdBlock *__restrict__ blocks,
                                                               int axis, const int *__restrict__ swr) {
  const GdBlock bk = gd_c(blocks)[blockIdx.x];
  GdJobC &J = gd_c(jobs)[bk.job];
  const int rnd = gd_c(swr)[bk.job];
  if (!gd_swap_active(J, rnd, axis)) return;
  const int lane = threadIdx.x & 63, wave = gd_uniform((int)(threadIdx.x >> 6));
  const int r = bk.rb + lane;
  const int Rp = J.Rp;
  bool lower;
  const int q = gd_swap_partner(J, r, axis, rnd, &lower);
  // a paired lane takes its partner's pose; the others keep their proposal
  // (prop of a paired lane is written below, never read here)
  GD_AS_GLOBAL float *prop = gd_g(J.prop);
  const GD_AS_GLOBAL float *mm = gd_g(J.mm);
  float m[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) m[i] = q >= 0 ? mm[(size_t)i * Rp + q] : prop[(size_t)i * Rp + r];
  if (wave == 0 && q >= 0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) prop[(size_t)i * Rp + r] = m[i];
    gd_g(J.swp_e)[r] = gd_g(J.en)[(size_t)9 * Rp + r];
  }
  float R[9];
  gd_rotation(m, R);
  const float t[3] = {m[0] + J.center[0], m[1] + J.center[1], m[2] + J.center[2]};
  gd_transform_atoms(J, r, R, t, wave, 4);
}

// Joint decision of each pair: with E the totals before the swap and T the
// totals at the partner's pose, both members accept iff
//   u < exp(mb_a (T_a - E_a) + mb_b (T_b - E_b)),
// a the lower-conformer member (both lanes evaluate the same expression), u
// from gd_swap_u01.  Accepted members take the pose and its ten terms.  Every
// lane then re-proposes the next MC step from its (possibly new) pose with the
// step's own draws and transforms it.
__global__ __launch_bounds__(256) void gd_swap_accept_kernel(const GdJob *__restrict__ jobs,
                                                              const GdBlock *__restrict__ blocks,
                                                              int axis, const int *__restrict__ swr,
                                                              int launch_step,
                                                              const int *__restrict__ step_base) {
  __shared__ float xf[12][64];
  const GdBlock bk = gd_c(blocks)[blockIdx.x];
  GdJobC &J = gd_c(jobs)[bk.job];
  const int rnd = gd_c(swr)[bk.job];
  if (!gd_swap_active(J, rnd, axis)) return;  // uniform over the block
  if (step_base) launch_step += *step_base;
  const int lane = threadIdx.x & 63, wave = gd_uniform((int)(threadIdx.x >> 6));
  const int r = bk.rb + lane;
  const int Rp = J.Rp;
  if (wave == 0) {
    GD_AS_GLOBAL float *prop = gd_g(J.prop), *mm = gd_g(J.mm), *en = gd_g(J.en);
    bool lower;
    const int q = gd_swap_partner(J, r, axis, rnd, &lower);
    float m[6];
    if (q >= 0) {
      const int a = lower ? r : q, b = lower ? q : r;
      const GD_AS_GLOBAL float *ev = gd_g(J.ev_terms), *e0 = gd_g(J.swp_e), *mb = gd_g(J.minus_beta);
      const GD_AS_GLOBAL int *tmpi = gd_g(J.tmpi);
      // products rounded separately (no fused multiply-add): for identical
      // conformers T_a - E_a = -(T_b - E_b) and x is exactly 0
      const float x = __fmul_rn(mb[tmpi[a]], ev[(size_t)a * 10 + 9] - e0[a]) +
                      __fmul_rn(mb[tmpi[b]], ev[(size_t)b * 10 + 9] - e0[b]);
      const float u = gd_swap_u01(((uint64_t)J.seed_hi << 32) | J.seed_lo, J.job_id, (uint32_t)a,
                                  (uint32_t)rnd, (uint32_t)axis);
      if (u < expf(x)) {
        float e[10];
#pragma unroll
        for (int i = 0; i < 10; ++i) e[i] = ev[(size_t)r * 10 + i];
#pragma unroll
        for (int i = 0; i < 10; ++i) en[(size_t)i * Rp + r] = e[i];
#pragma unroll
        for (int i = 0; i < 6; ++i) mm[(size_t)i * Rp + r] = prop[(size_t)i * Rp + r];
        GD_AS_GLOBAL float *best = gd_g(J.best);
        best[r] = fminf(best[r], e[9]);
        gd_g(J.nswp)[(size_t)axis * Rp + r] += 1;
        gd_g(J.acc)[r] = 1;
      }
      const int step = launch_step - J.step0;  // the MC step just finished
      float inc[6], un;
      gd_move(J, r, (uint32_t)step + 1u, inc, &un);
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const float v = inc[i] + mm[(size_t)i * Rp + r];
        prop[(size_t)i * Rp + r] = v;
        m[i] = v;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) m[i] = prop[(size_t)i * Rp + r];
    }
    float R[9];
    gd_rotation(m, R);
#pragma unroll
    for (int i = 0; i < 9; ++i) xf[i][lane] = R[i];
    xf[9][lane] = m[0] + J.center[0];
    xf[10][lane] = m[1] + J.center[1];
    xf[11][lane] = m[2] + J.center[2];
  }
  __syncthreads();
  float R[9], t[3];
#pragma unroll
  for (int i = 0; i < 9; ++i) R[i] = xf[i][lane];
  t[0] = xf[9][lane];
  t[1] = xf[10][lane];
  t[2] = xf[11][lane];
  gd_transform_atoms(J, r, R, t, wave, 4);
}

// ---------------------------------------------------------------------------
// Engine clock: one wave per SIMD of every CU spins on dependent FMAs for about
// `ticks` periods of the constant-rate counter (s_memrealtime) while reading
// the shader-clock counter (s_memtime); block 0's first wave records both
// deltas.  Every wave leaves once its own constant-rate delta passes `ticks`,
// or after 2^24 iterations (a few seconds) whatever the counters read.
__global__ __launch_bounds__(64) void gd_clock_kernel(long long ticks, float seed,
                                                       long long *__restrict__ out,
                                                       float *__restrict__ sink) {
  const long long w0 = wall_clock64(), c0 = clock64();
  float a = seed + (float)threadIdx.x, b = 1.0000001f;
  long long w1 = w0;
  for (int it = 0; it < (1 << 24) && w1 - w0 < ticks; ++it) {
#pragma unroll
    for (int i = 0; i < 64; ++i) a = __builtin_fmaf(a, b, 1e-7f);
    w1 = wall_clock64();
  }
  const long long c1 = clock64();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out[0] = c1 - c0;
    out[1] = w1 - w0;
  }
  if (a == 12345.0f) sink[0] = a;  // keeps the FMA chain
}

// ---------------------------------------------------------------------------
// Plain device copy (n16 16-byte vectors, grid-stride): the known-size
// reference against which HBM traffic counters are scaled.
__global__ __launch_bounds__(256) void gd_copy_kernel(const float4 *__restrict__ src,
                                                      float4 *__restrict__ dst, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256)
    dst[i] = src[i];
}
