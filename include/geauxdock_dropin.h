/*
 * geauxdock_dropin.h -- the reference GPU backend's own entry points, for
 * src/main_gpu/geauxdock.C (and main_mpi/client_gpu.C) to link against
 * libgeauxdock_amd.so in place of src/backend_gpu.
 *
 * The structs below restate, field for field, the reference's data model so
 * a Complex filled by the unchanged frontend (server1.C) is read correctly:
 *   Ligand/Protein/Psp/Kde/Mcs/Mcs_R/Mcs_ELL/Mcs_CSR/EnePara
 *                            src/common/geauxdock_soa.h:8-187
 *   Temp, ReplicaMC, McPara, Record, Complex, ParaT, ComplexSize
 *                            src/common/geauxdock.h:113-454
 * with the reference's capacities (src/common/size.h: MAXLIG 64, MAXPRO
 * 2048, MAX_CONF_LIG 64, MAX_CONF_PRT 20, MAXKDE 10240, MAX_MCS_ROW 512,
 * MAX_TMP 1024, MAX_REP 16384, STEPS_PER_DUMP 200) and 64-byte alignment
 * (geauxdock.h:29).  Larger complexes use the gd_* C ABI instead.
 *
 * Entry points replaced (C++ linkage, same names and argument meaning):
 *   void SetDevice();                                     prepair.h:4
 *   void DeviceAlloc(Complex**, Record**, gd_rng_state**) prepair.h:5
 *   void DeviceFree(Complex**, Record**, gd_rng_state**)  prepair.h:6
 *   void SetParaT(Complex*, ParaT**);                     prepair.h:7
 *   void CopyH2D(Complex*, Complex**, ParaT**);           prepair.h:8
 *   void InitCurand(gd_rng_state**);                      backend_gpu.h:4
 *   void Dock(Complex*, Record*, Complex**, Record**, ParaT**, gd_rng_state**);
 *                                                         backend_gpu.h:18-23
 * gd_rng_state (a Philox key + counter base) takes the place of cuRAND's
 * curandState: a caller declares its RNG pointers as gd_rng_state* (no CUDA
 * type name is defined here).  As in the reference, a failing HIP call prints
 * the error and exits.
 */
#ifndef GEAUXDOCK_DROPIN_H
#define GEAUXDOCK_DROPIN_H

#include <stdint.h>

#define GD_REF_MAXLIG 64
#define GD_REF_MAXPRO 2048
#define GD_REF_MAX_CONF_LIG 64
#define GD_REF_MAX_CONF_PRT 20
#define GD_REF_MAXKDE 10240
#define GD_REF_MAX_MCS_ROW 512
#define GD_REF_MAX_MCS_COL GD_REF_MAXLIG
#define GD_REF_MAX_TMP 1024
#define GD_REF_MAX_REP 16384
#define GD_REF_STEPS_PER_DUMP 200
#define GD_REF_MAXTP1 31
#define GD_REF_MAXTP2 24
#define GD_REF_MAXTP4 20
#define GD_REF_MAXWEI 10
#define GD_REF_MAXSTRINGLENG 128
#define GD_REF_MAX_STR_LENG 256
#define GD_REF_ALIGN __attribute__((aligned(64)))

/* A translation unit that already included the reference's geauxdock.h
 * (guard DOCK_H) uses the reference's own definitions: same bytes. */
#ifndef DOCK_H

struct GD_REF_ALIGN Ligand {
  float x1[GD_REF_MAXLIG], y1[GD_REF_MAXLIG], z1[GD_REF_MAXLIG];
  int t1[GD_REF_MAXLIG];
  float c1[GD_REF_MAXLIG];
  float x2[GD_REF_MAXLIG], y2[GD_REF_MAXLIG], z2[GD_REF_MAXLIG];
  int t2[GD_REF_MAXLIG];
  float c2[GD_REF_MAXLIG];
  float x3[GD_REF_MAXLIG], y3[GD_REF_MAXLIG], z3[GD_REF_MAXLIG];
  int t3[GD_REF_MAXLIG];
  float c3[GD_REF_MAXLIG];
  int kde_begin_idx[GD_REF_MAXLIG];
  int kde_end_idx[GD_REF_MAXLIG];
  float center[3];
  int lig_natom;
};

struct GD_REF_ALIGN Protein {
  int t[GD_REF_MAXPRO];
  float x[GD_REF_MAXPRO], y[GD_REF_MAXPRO], z[GD_REF_MAXPRO];
  int cdc[GD_REF_MAXPRO];
  float hpp[GD_REF_MAXPRO];
  float ele[GD_REF_MAXPRO];
  int c[GD_REF_MAXPRO];
  int seq3r[GD_REF_MAXPRO];
  float pocket_center[3];
  int prt_npoint;
};

struct GD_REF_ALIGN Psp {
  float psp[GD_REF_MAXLIG][GD_REF_MAXPRO];
};

struct GD_REF_ALIGN Kde {
  int t[GD_REF_MAXKDE];
  float x[GD_REF_MAXKDE], y[GD_REF_MAXKDE], z[GD_REF_MAXKDE];
  int kde_npoint;
};

struct GD_REF_ALIGN Mcs {
  float x[GD_REF_MAX_MCS_COL], y[GD_REF_MAX_MCS_COL], z[GD_REF_MAX_MCS_COL];
  int ncol;
  float tcc;
};

struct GD_REF_ALIGN Mcs_R {
  float x[GD_REF_MAX_MCS_COL][GD_REF_MAX_MCS_ROW];
  float y[GD_REF_MAX_MCS_COL][GD_REF_MAX_MCS_ROW];
  float z[GD_REF_MAX_MCS_COL][GD_REF_MAX_MCS_ROW];
  float tcc[GD_REF_MAX_MCS_ROW];
};

struct GD_REF_ALIGN Mcs_ELL {
  int i[GD_REF_MAX_MCS_ROW][GD_REF_MAX_MCS_COL];
  float x[GD_REF_MAX_MCS_ROW][GD_REF_MAX_MCS_COL];
  float y[GD_REF_MAX_MCS_ROW][GD_REF_MAX_MCS_COL];
  float z[GD_REF_MAX_MCS_ROW][GD_REF_MAX_MCS_COL];
  int ncol[GD_REF_MAX_MCS_ROW];
  float tcc[GD_REF_MAX_MCS_ROW];
};

struct GD_REF_ALIGN Mcs_CSR {
  int idx_col[GD_REF_MAX_MCS_COL * GD_REF_MAX_MCS_ROW];
  int idx_row[GD_REF_MAX_MCS_COL * GD_REF_MAX_MCS_ROW];
  float x[GD_REF_MAX_MCS_COL * GD_REF_MAX_MCS_ROW];
  float y[GD_REF_MAX_MCS_COL * GD_REF_MAX_MCS_ROW];
  float z[GD_REF_MAX_MCS_COL * GD_REF_MAX_MCS_ROW];
  float tcc[GD_REF_MAX_MCS_ROW];
  int npoint;
};

struct GD_REF_ALIGN EnePara {
  float p1a[GD_REF_MAXTP2][GD_REF_MAXTP1];
  float p2a[GD_REF_MAXTP2][GD_REF_MAXTP1];
  float lj0, lj1;
  float el1, el0, a1, b1;
  float pmf0[GD_REF_MAXTP2][GD_REF_MAXTP1];
  float pmf1[GD_REF_MAXTP2][GD_REF_MAXTP1];
  float hdb0[GD_REF_MAXTP2][GD_REF_MAXTP1];
  float hdb1[GD_REF_MAXTP2][GD_REF_MAXTP1];
  float hpp[GD_REF_MAXTP4];
  float hpl0[GD_REF_MAXTP2], hpl1[GD_REF_MAXTP2], hpl2[GD_REF_MAXTP2];
  float kde2, kde3;
  float a_para[GD_REF_MAXWEI], b_para[GD_REF_MAXWEI], w[GD_REF_MAXWEI];
};

struct ComplexSize {
  int n_lig, n_prt, n_tmp, n_rep;
  int lig_natom, prt_npoint, kde_npoint, mcs_nrow;
};

struct Temp {
  float t;
  float minus_beta;
  int order;
};

struct GD_REF_ALIGN ReplicaMC {
  int idx_rep, idx_lig, idx_prt, idx_tmp;
  float movematrix[6];
  float energy[GD_REF_MAXWEI];
  int step;
  int is_accept;
};

struct McPara {
  int steps_total;
  int steps_per_dump;
  int steps_per_exchange;
  float move_scale[6];
  char outputdir[GD_REF_MAXSTRINGLENG];
  char outputfile[GD_REF_MAXSTRINGLENG];
};

struct Record {
  ReplicaMC replica[GD_REF_STEPS_PER_DUMP];
  int next_entry;
};

struct Complex {
  Ligand lig[GD_REF_MAX_CONF_LIG];
  Protein prt[GD_REF_MAX_CONF_PRT];
  Psp psp;
  Kde kde;
  Mcs mcs[GD_REF_MAX_MCS_ROW];
  Mcs_R mcs_r;
  Mcs_ELL mcs_ell;
  Mcs_CSR mcs_csr;
  EnePara enepara;
  Temp temp[GD_REF_MAX_TMP];
  ReplicaMC replica[GD_REF_MAX_REP];
  ComplexSize size;
  McPara mcpara;
  char lig_file[GD_REF_MAX_STR_LENG];
  char prt_file[GD_REF_MAX_STR_LENG];
  char lhm_file[GD_REF_MAX_STR_LENG];
  char enepara_file[GD_REF_MAX_STR_LENG];
  char weight_file[GD_REF_MAX_STR_LENG];
  int rep_begin;
  int rep_end;
  int signal;
};

struct ParaT {
  int rep_begin;
  int rep_end;
  int n_rep;
  int record_sz;
};

#endif /* DOCK_H */

/* Counter-based RNG stream: Philox4x32-10 key and the job's stream id. */
struct gd_rng_state {
  uint64_t seed;
  uint32_t job_id;
  uint32_t reserved;
};

void SetDevice();
void DeviceAlloc(Complex **c, Record **r, gd_rng_state **s);
void DeviceFree(Complex **c, Record **r, gd_rng_state **s);
void SetParaT(Complex *ch, ParaT **pt);
void CopyH2D(Complex *ch, Complex **cd, ParaT **pt);
void InitCurand(gd_rng_state **s);
void Dock(Complex *ch, Record *rh, Complex **cd, Record **rd, ParaT **pt, gd_rng_state **s);

#endif
