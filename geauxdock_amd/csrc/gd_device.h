// gd_device.h -- device-resident layout of a docking job, shared by the
// kernels (gd_kernels.hip) and the host runtime (gd_runtime.hip).
//
// MI355X-first layout rules:
//  * replicas are the SIMD lanes: every per-replica array is SoA over a
//    replica index padded to a multiple of 64 (Rp), so a wave touches one
//    contiguous 256-byte line per field;
//  * everything shared by the replicas of a job (force-field rows, protein
//    point attributes, KDE anchors, MCS targets) is read through the scalar
//    data cache with wave-uniform addresses -- one fetch serves 64 lanes;
//  * only the protein coordinates depend on a lane (its protein conformer):
//    float4 [n_prt][P], <= a handful of distinct rows per wave;
//  * partial sums are [slot][Rp] so both the pair kernel (writer) and the
//    reduce kernel (reader) are coalesced across lanes.
#pragma once
#include <stdint.h>

#define GD_WAVE 64
#define GD_PTAB_W 12      // floats per (ligand type, point type) entry (9 used)
#define GD_PTAB_PT 32     // point-type stride of the table (31 used)
#define GD_REC_W 17       // floats per accept record: step, movematrix[6], energy[10]
#define GD_NLIN 6         // linear partial sums: vdw ele pmf psp hdb kde
#define GD_GROUP 8        // protein tiles (waves) per pair-kernel work group, summed in LDS
#define GD_RK 16          // reduction slices (work groups) per 64-replica block

// point flag bits (pattr.z as int)
#define GD_PF_TYPE 0x1f
#define GD_PF_CDC 0x100
#define GD_PF_CLS2 0x200

// finish-kernel modes
#define GD_MODE_INIT 0    // initial-energy pass: zero move, forced accept
#define GD_MODE_STEP 1    // MC step: Metropolis with the replica's draw
#define GD_MODE_EVAL 2    // energy only, results to the eval buffers

struct GdAtom {           // per ligand atom, type-sorted order
  int tidx;               // row of the job's parameter table
  float q;                // partial charge
  int kde_b, kde_e;       // KDE anchors of the atom's type [b, e)
  float kde_inv_n;        // 1 / (e - b), 0 if empty
  float hpl0, hpl1, hpl2; // hydrophobic restraint of the atom's type
};

struct GdJob {
  // sizes and tiling
  int L, P, K, M;
  int Ls;                 // atom stride of atoms / mcs: L padded to a multiple of 16 with
                          // dummy atoms whose every term is exactly zero
  int n_lig, n_prt, n_tmp, R, Rp;
  int nL, nP, ptile;      // ligand chunks, protein tiles (a multiple of GD_GROUP), points per tile
  int nG;                 // tile groups = nP / GD_GROUP: the partial-sum slots per chunk
  int pcls[5];            // points [pcls[c], pcls[c+1]) have class c = cdc + 2 * (class == 2)
  int lc;                 // ligand atoms per chunk (kernel template value)
  int n_types;            // rows of the parameter table
  // constant tables (device)
  const float4 *lig0;     // [n_lig][L] relative conformer coordinates
  const GdAtom *atoms;    // [Ls]
  const float4 *pcoord;   // [n_prt][P]
  const float4 *pattr;    // [P] (ele, hpp, flags-as-int, 0)
  const float *ptab;      // [n_types + 1][GD_PTAB_PT][GD_PTAB_W] (last block: padding atoms):
                          // 1/p1a, -p2a/p1a, negated PMF exponent B and C,
                          // pmf1, psp r^2 threshold (sign bit: H-bond active), H-bond hdb1, hdb0,
                          // far-field r^2 threshold
  const float *pspv;      // [n_types][P] psp value of class-2 points
  const float4 *kde;      // [K] sorted by type
  const float4 *mcs;      // [M][Ls] (x, y, z, valid)
  const float2 *mcsrow;   // [M] (tcc, n_valid)
  const float *minus_beta;// [n_tmp]
  // replica state (device, SoA over Rp)
  float *mm;              // [6][Rp] accepted movematrix
  float *prop;            // [6][Rp] proposed movematrix of the current step
  float *en;              // [10][Rp] accepted energies
  int *ligi, *prti, *tmpi;// [Rp] conformer / temperature indices
  int *acc;               // [Rp] last step accepted
  int *nacc, *nexc;       // [Rp] counters
  int *nswp;              // [2][Rp] accepted pose swaps (ligand axis, protein axis)
  float *swp_e;           // [Rp] total energy before a swap sub-step
  float *best;            // [Rp]
  int *own;               // [n_prt*n_lig][n_tmp] replica at each temperature
  float *coords;          // [L][3][Rp] transformed ligand of the current proposal
  float *part_lin;        // [GD_NLIN][nL*nG][Rp]
  float *part_hpc;        // [L][nG][Rp]
  float *part_mcs;        // [M][nL][Rp]
  float *fin_part;        // [Rp/64][GD_RK][8][64]: per-slice sums of the reduce kernel
  float *rec;             // [2][Rp][rec_cap][GD_REC_W] accept records (optional), a
                          // two-half ring: dump d fills half d mod 2
  int *rec_n;             // [2][Rp] entries used per half
  int rec_cap;
  float *ev_terms;        // [Rp][10] eval outputs
  float *ev_raw;          // [Rp][9]
  // scalars
  float lj0, lj1, el0, el1, a1, b1, inv_el1;
  float kde2l;            // kde2 * log2(e)
  float kde3_inv;
  float inv_L;
  float pmf_all;          // pmf1 summed over every (atom, point) pair of the job
  float center[3], pocket[3];
  float cp[3];            // center - pocket, rounded once from FP64 (DST term)
  float a_para[10], b_para[10], w[10];
  float move_scale[6];
  int move_mode;
  uint32_t seed_lo, seed_hi, job_id;
  int record;
  int xaxes;              // exchange axes (GD_EXCH_* bits)
  int step0;              // launch step at which this job's MC step 0 ran
  int steps_total;
  int *wq;                // work-queue counter of the persistent pair kernel
};

struct GdItem {           // one 4-wave work group of the pair kernel
  int job;
  int rb;                 // first replica of the 64-lane block
  int a;                  // ligand chunk
  int b;                  // protein tile group: tiles GD_GROUP*b .. GD_GROUP*b + GD_GROUP-1
};

struct GdBlock {          // one workgroup of the finish / transform kernels
  int job;
  int rb;
};

struct GdScale6 {         // move scales passed by value
  float v[6];
};

struct GdExch {           // one workgroup of the exchange kernel
  int job;
  int round;
};
