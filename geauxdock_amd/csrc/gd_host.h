// gd_host.h -- host-side complex ("job") built by the loaders or the synthetic
// generator; the data a gd_job_desc points into.
#pragma once
#include <string>
#include <vector>

#include "../../include/geauxdock_amd.h"

namespace gdr {
// the text gd_last_error() returns on this thread (defined with the runtime)
void set_error(const std::string &s);
}  // namespace gdr

struct gd_host_job {
  int n_lig = 0, n_prt = 0, n_tmp = 0;
  int L = 0, P = 0, K = 0, M = 0;
  // ligand [n_lig][L], type-sorted atom order
  std::vector<float> lig_x, lig_y, lig_z, lig_c;
  std::vector<int> lig_t, lig_perm;
  float lig_center[3] = {0, 0, 0};
  // protein [n_prt][P] coordinates, [P] attributes
  std::vector<float> prt_x, prt_y, prt_z, prt_hpp, prt_ele;
  std::vector<int> prt_t, prt_c, prt_cdc, prt_seq3r;
  float pocket[3] = {0, 0, 0};
  // psp [GD_NLT][psp_ncol]
  std::vector<float> psp;
  int psp_ncol = 0;
  // kde [K], sorted by type
  std::vector<int> kde_t;
  std::vector<float> kde_x, kde_y, kde_z;
  // mcs [M][L] in type-sorted atom order, tcc [M]
  std::vector<float> mcs_x, mcs_y, mcs_z, mcs_tcc;
  gd_enepara ene{};
  std::vector<float> minus_beta;
  // loader diagnostics
  int raw_conf = 0;                  // conformers listed in the SDF ensemble
  std::vector<float> conf_rmsd;      // superposed RMSD of every ensemble conformer to the native
  std::vector<int> mcs_ncol;         // matched atoms declared per MCS row
  std::string error;
};

namespace gdh {
// Reference code tables (frontend/data.C:238-587): -1 for unknown names,
// except ligand types which map unknown names to GD_BADTYPE as the reference does.
int lig_code(const std::string &s);
int res_code(const std::string &s);
int res_code_one(const std::string &s);
int pnt_code(const std::string &s);
const char *lig_name(int code);
float kabsch_rmsd(int n, const float *a, const float *b);
void set_temperature(int n, float floor_t, float ceil_t, float *minus_beta);
}  // namespace gdh
