// gd_host.cpp -- host side of the hot path: the reference frontend's loaders
// and typing rules, its SoA "Optimize" step, the temperature ladder, conformer
// RMSD, and the seeded synthetic-complex generator.  Plain C++17, no HIP.
//
// Parsing rules follow src/frontend/load.C (ligand :138-458, protein :637-1031,
// LHM :1054-1177, parameters :1189-1507) and the code tables of
// src/frontend/data.C:238-587; the SoA layout rules follow
// src/frontend/util_optimize_soa.C:36-540.  Deliberate differences, each a
// documented choice: every table is zero-initialised, unknown ligand atom
// types (BADKDE) reject the job instead of indexing out of bounds, sorts are
// stable, array sizes are dynamic, and errors are returned instead of exit().
#include "gd_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>

namespace gdh {

static const char *const kLigNames[GD_NLT] = {
    "Br",  "C.1", "C.2", "C.3", "C.ar", "C.cat", "Cl",    "F",   "I",   "N.1", "N.2", "N.3",
    "N.4", "N.am", "N.ar", "N.pl3", "O.2", "O.3", "O.co2", "P.3", "S.2", "S.3", "S.O", "S.O2"};
static const char *const kRes3[GD_NRES] = {"ALA", "ARG", "CYS", "PRO", "LEU", "ILE", "GLU",
                                           "HIS", "TRP", "THR", "MET", "ASP", "GLY", "PHE",
                                           "LYS", "GLN", "SER", "ASN", "TYR", "VAL"};
static const char kRes1[GD_NRES + 1] = "ARCPLIEHWTMDGFKQSNYV";
static const char *const kPnt[30] = {"CA", "PP", "A1", "C1", "D1", "E1", "E2", "F1", "F2", "H1",
                                     "H2", "I1", "K1", "K2", "L1", "M1", "M2", "N1", "P1", "Q1",
                                     "Q2", "R1", "R2", "S1", "T1", "V1", "W1", "W2", "Y1", "Y2"};

int lig_code(const std::string &s) {
  for (int i = 0; i < GD_NLT; ++i)
    if (s == kLigNames[i]) return i;
  return GD_BADTYPE;
}
const char *lig_name(int code) { return (code >= 0 && code < GD_NLT) ? kLigNames[code] : "?"; }
int res_code(const std::string &s) {
  for (int i = 0; i < GD_NRES; ++i)
    if (s == kRes3[i]) return i;
  return -1;
}
int res_code_one(const std::string &s) {
  if (s.size() != 1) return -1;
  for (int i = 0; i < GD_NRES; ++i)
    if (s[0] == kRes1[i]) return i;
  return -1;
}
int pnt_code(const std::string &s) {
  for (int i = 0; i < 30; ++i)
    if (s == kPnt[i]) return i;
  return -1;
}

// Side-chain effective points of each residue type: (point type, class) pairs,
// the expansion table of load.C:686-759.
struct SideChain {
  int n;
  int t[2];
  int c[2];
};
static const SideChain kSide[GD_NRES] = {
    {1, {2, 0}, {2, 0}},   {2, {21, 22}, {2, 3}}, {1, {3, 0}, {2, 0}},   {1, {18, 0}, {2, 0}},
    {1, {14, 0}, {2, 0}},  {1, {11, 0}, {2, 0}},  {2, {5, 6}, {2, 3}},   {2, {9, 10}, {2, 3}},
    {2, {26, 27}, {2, 3}}, {1, {24, 0}, {2, 0}},  {2, {15, 16}, {2, 3}}, {1, {4, 0}, {2, 0}},
    {0, {0, 0}, {0, 0}},   {2, {7, 8}, {2, 3}},   {2, {12, 13}, {2, 3}}, {2, {19, 20}, {2, 3}},
    {1, {23, 0}, {2, 0}},  {1, {17, 0}, {2, 0}},  {2, {28, 29}, {2, 3}}, {1, {25, 0}, {2, 0}}};

// Which PDB atoms average into a side-chain point (load.C:820-868):
// 1 = any non-backbone atom, 2 = non-backbone and not CB/CG, 3 = CB or CG.
static int side_selector(int t) {
  for (int r = 0; r < GD_NRES; ++r) {
    const SideChain &s = kSide[r];
    for (int k = 0; k < s.n; ++k)
      if (s.t[k] == t) return s.n == 1 ? 1 : (k == 0 ? 2 : 3);
  }
  return 0;
}

// Conformer RMSD after optimal rigid superposition of two centred point sets
// (Horn's quaternion method: the largest eigenvalue of the 4x4 key matrix of
// the correlation matrix gives the best-fit residual).  Replaces the Fortran
// u3b of src/frontend/rmsd.f:16, used at load.C:421-428.
float kabsch_rmsd(int n, const float *a, const float *b) {
  if (n <= 0) return 0.0f;
  double S[3][3] = {{0}}, e0 = 0;
  for (int i = 0; i < n; ++i) {
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) S[r][c] += (double)a[3 * i + r] * (double)b[3 * i + c];
      e0 += (double)a[3 * i + r] * a[3 * i + r] + (double)b[3 * i + r] * b[3 * i + r];
    }
  }
  double K[4][4] = {
      {S[0][0] + S[1][1] + S[2][2], S[1][2] - S[2][1], S[2][0] - S[0][2], S[0][1] - S[1][0]},
      {S[1][2] - S[2][1], S[0][0] - S[1][1] - S[2][2], S[0][1] + S[1][0], S[2][0] + S[0][2]},
      {S[2][0] - S[0][2], S[0][1] + S[1][0], -S[0][0] + S[1][1] - S[2][2], S[1][2] + S[2][1]},
      {S[0][1] - S[1][0], S[2][0] + S[0][2], S[1][2] + S[2][1], -S[0][0] - S[1][1] + S[2][2]}};
  // cyclic Jacobi eigenvalue iteration on the symmetric 4x4 matrix
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q) off += K[p][q] * K[p][q];
    if (off < 1e-22) break;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q) {
        if (std::fabs(K[p][q]) < 1e-300) continue;
        const double th = (K[q][q] - K[p][p]) / (2.0 * K[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 4; ++k) {
          const double kp = K[k][p], kq = K[k][q];
          K[k][p] = c * kp - s * kq;
          K[k][q] = s * kp + c * kq;
        }
        for (int k = 0; k < 4; ++k) {
          const double pk = K[p][k], qk = K[q][k];
          K[p][k] = c * pk - s * qk;
          K[q][k] = s * pk + c * qk;
        }
      }
  }
  double lmax = K[0][0];
  for (int i = 1; i < 4; ++i) lmax = std::max(lmax, K[i][i]);
  const double msd = std::max(0.0, e0 - 2.0 * lmax) / n;
  return (float)std::sqrt(msd);
}

// Geometric temperature ladder of util.C:325-363 (float arithmetic, as there).
void set_temperature(int n, float floor_t, float ceil_t, float *mb) {
  if (n == 1) {
    mb[0] = 0.0f - 1.0f / (1.0f * floor_t);
    return;
  }
  const float ratio = std::exp(std::log(ceil_t / floor_t) / (float)(n - 1));
  float a = floor_t;
  for (int i = 0; i < n; ++i) {
    mb[i] = 0.0f - 1.0f / (1.0f * a);
    a *= ratio;
  }
}
}  // namespace gdh

using namespace gdh;

// ---------------------------------------------------------------------------
// small text helpers
static bool read_lines(const char *path, std::vector<std::string> &out) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    out.push_back(line);
  }
  return true;
}
static std::vector<std::string> tokens(const std::string &s) {
  std::vector<std::string> t;
  std::istringstream is(s);
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}
static std::string sub(const std::string &s, size_t pos, size_t n) {
  return pos < s.size() ? s.substr(pos, n) : std::string();
}
// atof of a fixed-width field, converted to float as the reference does
static float fatof(const std::string &s) { return (float)std::atof(s.c_str()); }
// "istringstream >> float" parsing of the first 30 columns (load.C:1401-1404)
static float stream_float(const std::string &s) {
  std::istringstream os(s.substr(0, 30));
  float v = 0.0f;
  os >> v;
  return v;
}

// A loader's failure: the message goes to gd_last_error() and to stderr.
static gd_status loader_error(gd_status st, const std::string &msg) {
  gdr::set_error(msg);
  std::fprintf(stderr, "geauxdock_amd: %s\n", msg.c_str());
  return st;
}

// ---------------------------------------------------------------------------
// force field
extern "C" gd_status gd_load_enepara(const char *paras, const char *weight, const char *nor_a,
                                     const char *nor_b, gd_enepara *out) {
  if (!paras || !weight || !nor_a || !nor_b || !out) return GD_ERR_ARG;
  std::vector<std::string> lines;
  if (!read_lines(paras, lines))
    return loader_error(GD_ERR_IO, std::string("cannot open parameter file ") + paras);
  int lineno = 0;
  const auto bad = [&](gd_status st, const char *what) {
    return loader_error(st, std::string("parameter file ") + paras + " line " + std::to_string(lineno) +
                                ": " + what);
  };
  // raw tables (EnePara0 of geauxdock.h:279-295), zero-initialised, owned by
  // this call (concurrent loads share nothing)
  struct RawTables {
    float vdw[GD_NPT][GD_NLT][2], pmf[GD_NPT][GD_NLT][2], hdb[GD_NPT][GD_NLT][2];
  };
  std::unique_ptr<RawTables> raw(new RawTables());
  std::memset(raw.get(), 0, sizeof(RawTables));
  float(*vdw)[GD_NLT][2] = raw->vdw, (*pmf)[GD_NLT][2] = raw->pmf, (*hdb)[GD_NLT][2] = raw->hdb;
  float ele[GD_NELE] = {0}, hpp[GD_NRES] = {0}, hpl[GD_NLT][2] = {{0}};
  float lj[3] = {0, 0, 0}, el[2] = {0, 0}, kde = 0;
  float w[GD_NTERM] = {0}, a_para[GD_NTERM] = {0}, b_para[GD_NTERM] = {0};
  for (const std::string &ln : lines) {
    ++lineno;
    if (ln.size() <= 3) continue;
    const std::string key = ln.substr(0, 3);
    const std::vector<std::string> t = tokens(ln);
    if (key == "VDW" || key == "PMF" || key == "HDB") {
      if (t.size() < 5) return bad(GD_ERR_IO, "record needs a point type, a ligand type and two values");
      const int p = pnt_code(t[1]), l = lig_code(t[2]);
      if (p < 0 || l >= GD_NLT) return bad(GD_ERR_TYPE, "unknown point or ligand type");
      float(*tab)[GD_NLT][2] = key == "VDW" ? vdw : (key == "PMF" ? pmf : hdb);
      tab[p][l][0] = fatof(t[3]);
      tab[p][l][1] = fatof(t[4]);
    } else if (key == "PLJ") {
      for (int i = 0; i < 3 && i + 1 < (int)t.size(); ++i) lj[i] = fatof(t[i + 1]);
    } else if (key == "ELE") {
      if (t.size() < 3) return bad(GD_ERR_IO, "ELE record needs a point and a value");
      int idx;
      if (t[1].size() >= 2 && t[1][1] == 'C')
        idx = res_code_one(t[1].substr(0, 1)) + 30;  // residue CA entry (load.C:1245-1249)
      else
        idx = pnt_code(t[1]);
      if (idx < 0 || idx >= GD_NELE) return bad(GD_ERR_TYPE, "unknown ELE point");
      ele[idx] = fatof(t[2]);
    } else if (key == "PEL") {
      for (int i = 0; i < 2 && i + 1 < (int)t.size(); ++i) el[i] = fatof(t[i + 1]);
    } else if (key == "HPP") {
      if (t.size() < 3) return bad(GD_ERR_IO, "HPP record needs a residue and a value");
      const int r = res_code_one(t[1]);
      if (r < 0) return bad(GD_ERR_TYPE, "unknown HPP residue");
      hpp[r] = fatof(t[2]);
    } else if (key == "HPL") {
      if (t.size() < 4) return bad(GD_ERR_IO, "HPL record needs a ligand type and two values");
      const int l = lig_code(t[1]);
      if (l >= GD_NLT) return bad(GD_ERR_TYPE, "unknown HPL ligand type");
      hpl[l][0] = fatof(t[2]);
      hpl[l][1] = fatof(t[3]);
    } else if (key == "KDE") {
      if (t.size() >= 2) kde = fatof(t[1]);
    }
    // WEI is superseded by the weight file (loadWeight, load.C:1476-1506)
  }
  // hydrophobic scale min-max normalised to [-1, 1] (load.C:1351-1374)
  float lo = 1e6f, hi = -1e6f;
  for (int i = 0; i < GD_NRES; ++i) {
    lo = std::min(lo, hpp[i]);
    hi = std::max(hi, hpp[i]);
  }
  for (int i = 0; i < GD_NRES; ++i) {
    float v = (float)((double)((hpp[i] - lo) / (hi - lo)) * 2.0 - 1.0);
    hpp[i] = std::min(1.0f, std::max(-1.0f, v));
  }
  std::vector<std::string> wl, al, bl;
  const char *const files[3] = {weight, nor_a, nor_b};
  std::vector<std::string> *const dst[3] = {&wl, &al, &bl};
  for (int k = 0; k < 3; ++k)
    if (!read_lines(files[k], *dst[k]))
      return loader_error(GD_ERR_IO, std::string("cannot open parameter file ") + files[k]);
  for (size_t i = 0; i < wl.size() && i < GD_NTERM; ++i) w[i] = stream_float(wl[i]);
  for (size_t i = 0; i < al.size() && i < GD_NTERM; ++i) a_para[i] = stream_float(al[i]);
  for (size_t i = 0; i < bl.size() && i < GD_NTERM; ++i) b_para[i] = stream_float(bl[i]);

  // derived constants (OptimizeEnepara, util_optimize_soa.C:489-540)
  std::memset(out, 0, sizeof(*out));
  const float pi = 3.1415926535f;
  const float sqrt_2_pi = std::sqrt(2.0f * pi);
  for (int l = 0; l < GD_NLT; ++l)
    for (int p = 0; p < GD_NPT; ++p) {
      const float s = vdw[p][l][0] * lj[2];
      out->p1a[l][p] = 2.0f * vdw[p][l][1] * std::pow(s, 9.0f);
      out->p2a[l][p] = 3.0f * vdw[p][l][1] * std::pow(s, 6.0f);
      out->pmf0[l][p] = pmf[p][l][0];
      out->pmf1[l][p] = pmf[p][l][1];
      out->hdb0[l][p] = hdb[p][l][0];
      out->hdb1[l][p] = 1.0f / hdb[p][l][1];
    }
  out->lj0 = lj[0];
  out->lj1 = lj[1];
  out->el1 = el[1];
  out->el0 = el[0];
  out->a1 = 4.0f - 3.0f * el[0];
  out->b1 = 2.0f * el[0] - 3.0f;
  for (int i = 0; i < GD_NRES; ++i) out->hpp[i] = hpp[i];
  for (int l = 0; l < GD_NLT; ++l) {
    out->hpl0[l] = hpl[l][0];
    out->hpl1[l] = hpl[l][1];
    out->hpl2[l] = std::log(1.0f / (out->hpl1[l] * sqrt_2_pi));
  }
  out->kde2 = -0.5f / (kde * kde);
  out->kde3 = std::pow(kde * sqrt_2_pi, 3.0f);
  for (int i = 0; i < GD_NTERM; ++i) {
    out->w[i] = w[i];
    out->a_para[i] = a_para[i];
    out->b_para[i] = b_para[i];
  }
  for (int i = 0; i < GD_NELE; ++i) out->ele[i] = ele[i];
  return GD_OK;
}

// ---------------------------------------------------------------------------
// ligand (.sdf with OpenBabel property blocks and an ENSEMBLE_COORDS block)
struct Lig0 {
  int natom = 0;
  std::vector<int> t;
  std::vector<float> c;
  std::vector<float> xyz;                // native, relative to its centroid [natom][3]
  std::vector<std::vector<float>> conf;  // accepted ensemble conformers, centred
  float center[3] = {0, 0, 0};
  int raw_conf = 0;
  std::vector<float> rmsd;
};

static gd_status load_ligand(const char *path, Lig0 &lig, std::string &err) {
  std::vector<std::string> ln;
  if (!read_lines(path, ln)) {
    err = std::string("cannot open ligand file ") + path;
    return GD_ERR_IO;
  }
  size_t end = 0;
  while (end < ln.size() && ln[end] != "$$$$") ++end;
  if (end >= ln.size() || end <= 10) {
    err = "ligand file has no complete record";
    return GD_ERR_IO;
  }
  const int natom = std::atoi(sub(ln[3], 0, 3).c_str());
  const int nbond = std::atoi(sub(ln[3], 3, 3).c_str());
  if (natom <= 0 || 4 + natom + nbond > (int)end) {
    err = "bad SDF counts line";
    return GD_ERR_IO;
  }
  lig.natom = natom;
  lig.t.assign(natom, GD_BADTYPE);
  lig.c.assign(natom, 0.0f);
  std::vector<std::vector<float>> ens;
  int have_types = 0;
  for (size_t i = 4 + natom + nbond; i + 1 < end; ++i) {
    const std::string &s = ln[i];
    if (s.find("OB_ATOM_TYPES") != std::string::npos) {
      const auto t = tokens(ln[i + 1]);
      for (int k = 0; k < natom && k < (int)t.size(); ++k) lig.t[k] = lig_code(t[k]);
      have_types = (int)t.size();
    } else if (s.find("OB_ATOMIC_CHARGES") != std::string::npos) {
      const auto t = tokens(ln[i + 1]);
      for (int k = 0; k < natom && k < (int)t.size(); ++k) lig.c[k] = fatof(t[k]);
    } else if (s.find("ENSEMBLE_COORDS") != std::string::npos) {
      while (i + 1 < end && !ln[i + 1].empty() && (int)ens.size() < 64) {
        ++i;
        const auto t = tokens(ln[i]);
        std::vector<float> v(3 * natom, 0.0f);
        for (int k = 0; k < 3 * natom && k < (int)t.size(); ++k) v[k] = fatof(t[k]);
        ens.push_back(v);
      }
    }
  }
  if (have_types < natom) {
    err = "missing OB_ATOM_TYPES";
    return GD_ERR_IO;
  }
  for (int k = 0; k < natom; ++k)
    if (lig.t[k] >= GD_NLT) {
      err = "unknown SYBYL atom type (BADKDE) in ligand";
      return GD_ERR_TYPE;
    }
  // native pose: centroid accumulated in float in file order (load.C:296-320)
  std::vector<float> nat(3 * natom);
  for (int k = 0; k < natom; ++k) {
    const std::string &a = ln[4 + k];
    nat[3 * k + 0] = fatof(sub(a, 0, 10));
    nat[3 * k + 1] = fatof(sub(a, 10, 10));
    nat[3 * k + 2] = fatof(sub(a, 20, 10));
    for (int d = 0; d < 3; ++d) lig.center[d] += nat[3 * k + d];
  }
  for (int d = 0; d < 3; ++d) lig.center[d] /= (float)natom;
  for (int k = 0; k < natom; ++k)
    for (int d = 0; d < 3; ++d) nat[3 * k + d] -= lig.center[d];
  lig.xyz = nat;
  lig.raw_conf = (int)ens.size();
  // ensemble conformers: centre each, keep those with RMSD > 0.1 A to the
  // native after superposition (load.C:362-456); at most 64 conformers total
  lig.conf.clear();
  for (auto &m : ens) {
    float cen[3] = {0, 0, 0};
    for (int k = 0; k < natom; ++k)
      for (int d = 0; d < 3; ++d) cen[d] += m[3 * k + d];
    for (int d = 0; d < 3; ++d) cen[d] /= (float)natom;
    for (int k = 0; k < natom; ++k)
      for (int d = 0; d < 3; ++d) m[3 * k + d] -= cen[d];
    const float r = kabsch_rmsd(natom, m.data(), nat.data());
    lig.rmsd.push_back(r);
    if (r > 0.1f) lig.conf.push_back(m);
    if ((int)lig.conf.size() + 1 >= 64) break;
  }
  return GD_OK;
}

// ---------------------------------------------------------------------------
// eFindSite restraint file: KDE points, PSP table, MCS rows (load.C:1054-1177)
struct Lhm0 {
  std::vector<int> kt;
  std::vector<float> kx, ky, kz;
  std::map<std::pair<int, int>, float> psp;  // (residue serial, ligand type) -> value
  std::vector<std::vector<float>> mx, my, mz;  // [row][natom] file atom order
  std::vector<float> tcc;
  std::vector<int> ncol;
};

static gd_status load_lhm(const char *path, const std::string &lig_id, int natom, Lhm0 &h,
                          std::string &err) {
  std::vector<std::string> ln;
  if (!read_lines(path, ln)) {
    err = std::string("cannot open restraint file ") + path;
    return GD_ERR_IO;
  }
  int lineno = 0;
  const auto bad = [&](const char *what) {
    err = std::string("restraint file ") + path + " line " + std::to_string(lineno) + ": " + what;
    return GD_ERR_IO;
  };
  for (const std::string &s : ln) {
    ++lineno;
    if (s.size() <= 3) continue;
    const std::string key = s.substr(0, 3);
    if (key == "KDE") {
      const auto t = tokens(s);
      if (t.size() < 5) return bad("KDE record needs a type and three coordinates");
      h.kt.push_back(lig_code(t[1]));
      h.kx.push_back(fatof(t[2]));
      h.ky.push_back(fatof(t[3]));
      h.kz.push_back(fatof(t[4]));
    } else if (key == "PSP") {
      const auto t = tokens(s);
      if (t.size() < 4) return bad("PSP record needs a residue, a type and a value");
      const int res = std::atoi(t[1].c_str()), lt = lig_code(t[2]);
      if (res < 0) return bad("PSP residue number is negative");
      if (lt < GD_NLT) h.psp[{res, lt}] = fatof(t[3]);
    } else if (key == "MCS") {
      const auto t = tokens(s);
      std::vector<float> x(natom, GD_MCS_INVALID), y(natom, GD_MCS_INVALID),
          z(natom, GD_MCS_INVALID);
      float tcc = 0.0f;
      int ncol = 0;
      if (t.size() >= 4 && t[1] == lig_id) {
        tcc = fatof(t[2]);
        ncol = std::atoi(t[3].c_str());
        for (int ia = 0; ia < ncol; ++ia) {
          if ((size_t)(ia * 4 + 7) >= t.size()) return bad("MCS record has fewer columns than it declares");
          const int loc = std::atoi(t[ia * 4 + 4].c_str()) - 1;
          if (loc < 0 || loc >= 63) return bad("MCS position out of bounds");
          if (loc >= natom) continue;
          x[loc] = fatof(t[ia * 4 + 5]);
          y[loc] = fatof(t[ia * 4 + 6]);
          z[loc] = fatof(t[ia * 4 + 7]);
        }
      }
      // rows of other ligand ids still count (load.C:1162-1163)
      h.mx.push_back(x);
      h.my.push_back(y);
      h.mz.push_back(z);
      h.tcc.push_back(tcc);
      h.ncol.push_back(ncol);
    }
  }
  return GD_OK;
}

// ---------------------------------------------------------------------------
// protein (.pdb, coarse-grained to 1-4 effective points per residue)
struct Prt0 {
  int npoint = 0, pnr = 0, nconf = 0;
  std::vector<int> r, t, d, c;
  std::vector<int> seq3;                 // residue serial of each CA, [pnr]
  std::vector<std::vector<float>> x, y, z;  // [nconf][npoint]
};

static gd_status load_protein(const char *path, Prt0 &p, std::string &err) {
  std::vector<std::string> ln;
  if (!read_lines(path, ln)) {
    err = std::string("cannot open protein file ") + path;
    return GD_ERR_IO;
  }
  auto is_atom = [](const std::string &s) { return s.size() > 53 && s.compare(0, 6, "ATOM  ") == 0; };
  auto is_endmdl = [](const std::string &s) { return s.compare(0, 6, "ENDMDL") == 0; };
  // point list from the first model (load.C:686-766)
  for (const std::string &s : ln) {
    if (is_atom(s)) {
      if (s.substr(12, 4) != " CA ") continue;
      const int res = res_code(s.substr(17, 3));
      if (res < 0) {
        err = "unknown residue " + s.substr(17, 3);
        return GD_ERR_TYPE;
      }
      const int serial = std::atoi(s.substr(22, 4).c_str());
      const int i = p.pnr;
      p.seq3.push_back(serial);
      auto push = [&](int r, int t, int c) {
        p.r.push_back(r);
        p.t.push_back(t);
        p.d.push_back(res);
        p.c.push_back(c);
      };
      push(i, 0, 0);
      if (i > 1) push(i - 1, 1, 1);
      const SideChain &sc = kSide[res];
      for (int k = 0; k < sc.n; ++k) push(i, sc.t[k], sc.c[k]);
      ++p.pnr;
    } else if (s.size() <= 53 && is_endmdl(s)) {
      break;
    }
  }
  p.npoint = (int)p.t.size();
  // coordinates: per point, average its atoms; every ENDMDL closes a conformer
  for (const std::string &s : ln)
    if (s.size() <= 53 && is_endmdl(s)) ++p.nconf;
  if (p.nconf == 0) {
    err = "protein file has no ENDMDL record";
    return GD_ERR_IO;
  }
  p.x.assign(p.nconf, std::vector<float>(p.npoint, 0.0f));
  p.y = p.x;
  p.z = p.x;
  // pre-parse atom lines once
  struct A {
    int line, serial;
    std::string name;
    float x, y, z;
  };
  std::vector<A> atoms;
  std::vector<int> endmdl_after;  // atom count seen before each ENDMDL
  for (int i = 0; i < (int)ln.size(); ++i) {
    const std::string &s = ln[i];
    if (is_atom(s)) {
      atoms.push_back({i, std::atoi(s.substr(22, 4).c_str()), s.substr(12, 4),
                       fatof(s.substr(30, 8)), fatof(s.substr(38, 8)), fatof(s.substr(46, 8))});
    } else if (s.size() <= 53 && is_endmdl(s)) {
      endmdl_after.push_back((int)atoms.size());
    }
  }
  for (int q = 0; q < p.npoint; ++q) {
    const int rq = p.r[q], tq = p.t[q];
    const int serial = p.seq3[rq];
    const int sel = side_selector(tq);
    const bool pp = (tq == 1 && rq > 0 && rq < p.pnr - 1);
    const int prev = (tq == 1 && rq > 0) ? p.seq3[rq - 1] : -1;
    float sx = 0, sy = 0, sz = 0, sn = 0;
    int conf = 0, a = 0;
    for (int m = 0; m < (int)endmdl_after.size(); ++m) {
      for (; a < endmdl_after[m]; ++a) {
        const A &at = atoms[a];
        const std::string &nm = at.name;
        bool take = false;
        if (at.serial == serial) {
          const bool bb = nm == " N  " || nm == " CA " || nm == " C  " || nm == " O  ";
          const bool cbcg = nm == " CB " || nm == " CG ";
          if (tq == 0)
            take = nm == " CA ";
          else if (sel == 1)
            take = !bb;
          else if (sel == 2)
            take = !bb && !cbcg;
          else if (sel == 3)
            take = cbcg;
        }
        if (pp && ((at.serial == serial && nm == " N  ") ||
                   (at.serial == prev && (nm == " C  " || nm == " O  "))))
          take = true;
        if (take) {
          sx += at.x;
          sy += at.y;
          sz += at.z;
          sn += 1.0f;
        }
      }
      if (sn > 0.0f) {
        p.x[conf][q] = sx / sn;
        p.y[conf][q] = sy / sn;
        p.z[conf][q] = sz / sn;
      }
      sx = sy = sz = sn = 0;
      ++conf;
    }
  }
  return GD_OK;
}

// ---------------------------------------------------------------------------
// assembly of one complex (server1.C:54-217 + Optimize*)
static void finish_ligand(gd_host_job &h, const std::vector<int> &t_file,
                          const std::vector<float> &c_file,
                          const std::vector<std::vector<float>> &confs /* file order, [n][3L] */) {
  const int L = (int)t_file.size();
  h.L = L;
  h.n_lig = (int)confs.size();
  // stable sort of atoms by SYBYL code (x3/t3/c3 order, util_optimize_soa.C:112-126)
  std::vector<int> perm(L);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return t_file[a] < t_file[b]; });
  h.lig_perm = perm;
  h.lig_t.resize(L);
  h.lig_c.resize(L);
  for (int j = 0; j < L; ++j) {
    h.lig_t[j] = t_file[perm[j]];
    h.lig_c[j] = c_file[perm[j]];
  }
  h.lig_x.resize((size_t)h.n_lig * L);
  h.lig_y.resize((size_t)h.n_lig * L);
  h.lig_z.resize((size_t)h.n_lig * L);
  for (int i = 0; i < h.n_lig; ++i)
    for (int j = 0; j < L; ++j) {
      h.lig_x[(size_t)i * L + j] = confs[i][3 * perm[j] + 0];
      h.lig_y[(size_t)i * L + j] = confs[i][3 * perm[j] + 1];
      h.lig_z[(size_t)i * L + j] = confs[i][3 * perm[j] + 2];
    }
}

static void finish_kde(gd_host_job &h, const std::vector<int> &kt, const std::vector<float> &kx,
                       const std::vector<float> &ky, const std::vector<float> &kz) {
  const int K = (int)kt.size();
  std::vector<int> o(K);
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return kt[a] < kt[b]; });
  h.K = K;
  h.kde_t.resize(K);
  h.kde_x.resize(K);
  h.kde_y.resize(K);
  h.kde_z.resize(K);
  for (int i = 0; i < K; ++i) {
    h.kde_t[i] = kt[o[i]];
    h.kde_x[i] = kx[o[i]];
    h.kde_y[i] = ky[o[i]];
    h.kde_z[i] = kz[o[i]];
  }
}

// protein SoA: class-sorted points, derived attributes (util_optimize_soa.C:198-269)
static void finish_protein(gd_host_job &h, const Prt0 &p, const gd_enepara &ene) {
  const int P = p.npoint;
  h.P = P;
  h.n_prt = p.nconf;
  std::vector<int> o(P);
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return p.c[a] < p.c[b]; });
  h.prt_t.resize(P);
  h.prt_c.resize(P);
  h.prt_cdc.resize(P);
  h.prt_seq3r.resize(P);
  h.prt_hpp.resize(P);
  h.prt_ele.resize(P);
  for (int j = 0; j < P; ++j) {
    const int q = o[j];
    const int t = p.t[q], d = p.d[q], c = p.c[q];
    h.prt_t[j] = t;
    h.prt_c[j] = c;
    h.prt_cdc[j] = (c == 0 && d == 12) || c == 2;
    h.prt_seq3r[j] = p.seq3[p.r[q]];
    h.prt_ele[j] = ene.ele[t == 0 ? d + 30 : t];
    h.prt_hpp[j] = ene.hpp[d];
  }
  h.prt_x.resize((size_t)h.n_prt * P);
  h.prt_y.resize((size_t)h.n_prt * P);
  h.prt_z.resize((size_t)h.n_prt * P);
  for (int i = 0; i < h.n_prt; ++i)
    for (int j = 0; j < P; ++j) {
      h.prt_x[(size_t)i * P + j] = p.x[i][o[j]];
      h.prt_y[(size_t)i * P + j] = p.y[i][o[j]];
      h.prt_z[(size_t)i * P + j] = p.z[i][o[j]];
    }
}

static void set_psp(gd_host_job &h, const std::map<std::pair<int, int>, float> &psp) {
  int maxres = 0;
  for (int s : h.prt_seq3r) maxres = std::max(maxres, s);
  for (auto &kv : psp) maxres = std::max(maxres, kv.first.first);
  h.psp_ncol = maxres + 1;
  h.psp.assign((size_t)GD_NLT * h.psp_ncol, 0.0f);
  for (auto &kv : psp) h.psp[(size_t)kv.first.second * h.psp_ncol + kv.first.first] = kv.second;
}

extern "C" gd_status gd_load_complex(const char *sdf, const char *ff, const char *pdb,
                                     const char *ligand_id, const gd_enepara *ene, int n_tmp,
                                     float floor_temp, float ceiling_temp, gd_host_job **out) {
  if (!sdf || !ff || !pdb || !ligand_id || !ene || !out || n_tmp < 1) return GD_ERR_ARG;
  *out = nullptr;
  gd_host_job *h = new gd_host_job();
  std::string err;
  Lig0 lig;
  gd_status st = load_ligand(sdf, lig, err);
  if (st != GD_OK) {
    delete h;
    return loader_error(st, err);
  }
  Lhm0 lhm;
  st = load_lhm(ff, ligand_id, lig.natom, lhm, err);
  if (st != GD_OK) {
    delete h;
    return loader_error(st, err);
  }
  Prt0 prt;
  st = load_protein(pdb, prt, err);
  if (st != GD_OK) {
    delete h;
    return loader_error(st, err);
  }
  h->ene = *ene;
  std::vector<std::vector<float>> confs;
  confs.push_back(lig.xyz);
  for (auto &c : lig.conf) confs.push_back(c);
  finish_ligand(*h, lig.t, lig.c, confs);
  for (int d = 0; d < 3; ++d) {
    h->lig_center[d] = lig.center[d];
    h->pocket[d] = lig.center[d];  // SetPocketCenter with the native centroid (server1.C:100-103)
  }
  h->raw_conf = lig.raw_conf;
  h->conf_rmsd = lig.rmsd;
  finish_kde(*h, lhm.kt, lhm.kx, lhm.ky, lhm.kz);
  finish_protein(*h, prt, *ene);
  set_psp(*h, lhm.psp);
  // MCS rows, permuted into the type-sorted atom order
  h->M = (int)lhm.tcc.size();
  const int L = h->L;
  h->mcs_x.resize((size_t)h->M * L);
  h->mcs_y.resize((size_t)h->M * L);
  h->mcs_z.resize((size_t)h->M * L);
  for (int m = 0; m < h->M; ++m)
    for (int j = 0; j < L; ++j) {
      const int f = h->lig_perm[j];
      h->mcs_x[(size_t)m * L + j] = lhm.mx[m][f];
      h->mcs_y[(size_t)m * L + j] = lhm.my[m][f];
      h->mcs_z[(size_t)m * L + j] = lhm.mz[m][f];
    }
  h->mcs_tcc = lhm.tcc;
  h->mcs_ncol = lhm.ncol;
  h->n_tmp = n_tmp;
  h->minus_beta.resize(n_tmp);
  set_temperature(n_tmp, floor_temp, ceiling_temp, h->minus_beta.data());
  *out = h;
  return GD_OK;
}

// ---------------------------------------------------------------------------
// synthetic complexes (seeded; sizes of the benchmark workloads)
//
// Protein: residues of uniformly random type expanded into effective points by
// the loader's own table until exactly P points exist; CA trace = 3.8 A random
// walk inside a box of 40 A * (P/1200)^(1/3) that avoids a pocket sphere;
// side-chain and peptide points sit 1.5-2.5 A off the trace.  Per-point
// charges ~N(0, 0.3) shifted to a net zero; conformers are 0.3 A Gaussian
// jitters of conformer 0.  Ligand: random bonded tree, 1.5 A bonds, angles
// 109.5 +- 10 deg, C/N/O 70/15/15 % mapped to SYBYL C.3/C.ar/C.2, N.am/N.ar/
// N.3, O.2/O.3/O.co2; charges ~N(0, 0.2); conformers perturb torsions by up
// to +-30 deg.  Restraints: 120 KDE points per atom of the atom's type at
// N(native, 1 A); 64 MCS rows matching ~50 % of atoms at N(native, 1 A) with
// tcc ~ U(0.1, 1); 32 PSP entries (nearest residues, random ligand type
// present) ~ N(0, 1).
namespace {
struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t s) : g(s) {}
  double u() { return std::uniform_real_distribution<double>(0.0, 1.0)(g); }
  double n() { return std::normal_distribution<double>(0.0, 1.0)(g); }
  int i(int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(g); }
};
struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(V3 a) { return std::sqrt(dot(a, a)); }
V3 unit(V3 a) { return (1.0 / norm(a)) * a; }
V3 rand_dir(Rng &r) {
  for (;;) {
    V3 v{r.n(), r.n(), r.n()};
    const double l = norm(v);
    if (l > 1e-6) return (1.0 / l) * v;
  }
}
// rotate p about the axis through o with unit direction k by angle a (Rodrigues)
V3 rot_about(V3 p, V3 o, V3 k, double a) {
  V3 v = p - o;
  const double c = std::cos(a), s = std::sin(a);
  V3 r = c * v + s * cross(k, v) + (dot(k, v) * (1 - c)) * k;
  return o + r;
}
}  // namespace

// The synthetic complex of the benchmark workloads.  rl draws the ligand and
// its restraints (KDE, MCS, PSP), rp the protein and its charges; with one
// generator for both the draws interleave in that order.  pocket_r > 0 fixes
// the radius of the protein's empty pocket (one protein for many ligands),
// else it is the ligand's radius + 2.5 A.
static gd_status synth(int P, int L, int n_lig, int n_prt, int n_tmp, float floor_temp,
                       float ceiling_temp, Rng &rl, Rng &rp, double pocket_r_fixed,
                       const gd_enepara *ene, gd_host_job **out) {
  if (!ene || !out || P < 8 || L < 2 || n_lig < 1 || n_prt < 1 || n_tmp < 1) return GD_ERR_ARG;
  *out = nullptr;
  Rng &rng = rl;
  gd_host_job *h = new gd_host_job();
  h->ene = *ene;
  const double box = 40.0 * std::cbrt(P / 1200.0);
  const V3 center{box / 2, box / 2, box / 2};

  // ---- ligand: random tree around the origin
  std::vector<V3> la;
  std::vector<int> parent, ndeg;
  std::vector<int> lt(L);
  la.push_back({0, 0, 0});
  parent.push_back(-1);
  ndeg.push_back(0);
  const double ang0 = 109.5 * M_PI / 180.0;
  while ((int)la.size() < L) {
    const int a = rng.i(0, (int)la.size() - 1);
    if (ndeg[a] >= 3) continue;
    V3 d;
    if (parent[a] < 0) {
      d = rand_dir(rng);
    } else {
      // bond at 109.5 +- 10 deg from the parent bond, random azimuth
      const V3 b = unit(la[parent[a]] - la[a]);
      const double th = ang0 + (rng.u() * 2 - 1) * 10.0 * M_PI / 180.0;
      V3 perp = unit(cross(b, rand_dir(rng)));
      d = std::cos(th) * b + std::sin(th) * perp;
    }
    const V3 p = la[a] + 1.5 * d;
    bool clash = false;
    for (size_t k = 0; k < la.size(); ++k)
      if ((int)k != a && norm(la[k] - p) < 2.2) clash = true;
    if (clash) continue;
    la.push_back(p);
    parent.push_back(a);
    ndeg.push_back(1);
    ++ndeg[a];
  }
  static const int kC[3] = {3, 4, 2}, kN[3] = {13, 14, 11}, kO[3] = {16, 17, 18};
  std::vector<float> lc(L);
  for (int j = 0; j < L; ++j) {
    const double e = rng.u();
    const int k = rng.i(0, 2);
    lt[j] = e < 0.70 ? kC[k] : (e < 0.85 ? kN[k] : kO[k]);
    lc[j] = (float)(0.2 * rng.n());
  }
  // conformers: torsion moves of subtrees
  std::vector<std::vector<V3>> lconf;
  lconf.push_back(la);
  for (int c = 1; c < n_lig; ++c) {
    std::vector<V3> x = la;
    for (int k = 1; k < L; ++k) {
      if (parent[k] < 0) continue;
      const int pa = parent[k];
      const double a = (rng.u() * 2 - 1) * 30.0 * M_PI / 180.0;
      const V3 axis = unit(x[k] - x[pa]);
      // subtree of k
      std::vector<char> in(L, 0);
      in[k] = 1;
      for (int q = k + 1; q < L; ++q)
        if (parent[q] >= 0 && in[parent[q]]) in[q] = 1;
      for (int q = 0; q < L; ++q)
        if (in[q] && q != k) x[q] = rot_about(x[q], x[k], axis, a);
    }
    lconf.push_back(x);
  }
  // centre every conformer on its own centroid; native centroid = pocket centre
  std::vector<std::vector<float>> confs;
  for (auto &x : lconf) {
    V3 m{0, 0, 0};
    for (auto &p : x) m = m + p;
    m = (1.0 / L) * m;
    std::vector<float> v(3 * L);
    for (int j = 0; j < L; ++j) {
      v[3 * j + 0] = (float)(x[j].x - m.x);
      v[3 * j + 1] = (float)(x[j].y - m.y);
      v[3 * j + 2] = (float)(x[j].z - m.z);
    }
    confs.push_back(v);
  }
  double lig_radius = 0;
  for (int j = 0; j < L; ++j)
    lig_radius = std::max(lig_radius, norm(V3{confs[0][3 * j], confs[0][3 * j + 1], confs[0][3 * j + 2]}));
  finish_ligand(*h, lt, lc, confs);
  for (int d = 0; d < 3; ++d) {
    h->lig_center[d] = (float)(d == 0 ? center.x : d == 1 ? center.y : center.z);
    h->pocket[d] = h->lig_center[d];
  }
  std::vector<V3> native(L);
  for (int j = 0; j < L; ++j)
    native[j] = center + V3{confs[0][3 * j], confs[0][3 * j + 1], confs[0][3 * j + 2]};

  // ---- protein: residue chains by random walk, expanded into points
  Prt0 pr;
  const double pocket_r = pocket_r_fixed > 0.0 ? pocket_r_fixed : lig_radius + 2.5;
  // Points go in the box outside the pocket sphere.  A box whose corners lie
  // within the sphere has no room at all; one whose room is a sliver would
  // take unbounded rejection sampling, so chain starts get a fixed budget.
  const auto no_room = [&]() {
    char msg[160];
    std::snprintf(msg, sizeof msg,
                  "synthetic protein of %d points (box %.1f A) has no room outside a pocket of "
                  "radius %.1f A",
                  P, box, pocket_r);
    delete h;
    return loader_error(GD_ERR_ARG, msg);
  };
  if (pocket_r >= 0.5 * std::sqrt(3.0) * box) return no_room();
  std::vector<V3> cas;  // CA of each residue
  {
  Rng &rng = rp;
  auto inside = [&](V3 p) {
    return p.x > 0 && p.y > 0 && p.z > 0 && p.x < box && p.y < box && p.z < box &&
           norm(p - center) > pocket_r;
  };
  V3 ca{0, 0, 0};
  bool new_chain = true;
  int serial = 0;
  std::vector<V3> pts;
  while ((int)pr.t.size() < P) {
    if (new_chain) {
      int tries = 0;
      do {
        ca = V3{rng.u() * box, rng.u() * box, rng.u() * box};
      } while (!inside(ca) && ++tries < (1 << 20));
      if (!inside(ca)) return no_room();
      new_chain = false;
    } else {
      V3 nx;
      int tries = 0;
      do {
        nx = ca + 3.8 * rand_dir(rng);
      } while (!inside(nx) && ++tries < 64);
      if (tries >= 64) {
        new_chain = true;
        continue;
      }
      ca = nx;
    }
    const int res = rng.i(0, GD_NRES - 1);
    const int i = pr.pnr;
    ++serial;
    pr.seq3.push_back(serial);
    cas.push_back(ca);
    auto push = [&](int r, int t, int c, V3 x) {
      if ((int)pr.t.size() >= P) return;
      pr.r.push_back(r);
      pr.t.push_back(t);
      pr.d.push_back(res);
      pr.c.push_back(c);
      pts.push_back(x);
    };
    push(i, 0, 0, ca);
    if (i > 1) push(i - 1, 1, 1, 0.5 * (ca + cas[i - 1]) + 0.8 * rand_dir(rng));
    const SideChain &sc = kSide[res];
    V3 sd = rand_dir(rng);
    for (int k = 0; k < sc.n; ++k) push(i, sc.t[k], sc.c[k], ca + (1.5 + 1.0 * k) * sd + 0.3 * rand_dir(rng));
    ++pr.pnr;
  }
  pr.npoint = P;
  pr.nconf = n_prt;
  pr.x.assign(n_prt, std::vector<float>(P));
  pr.y = pr.x;
  pr.z = pr.x;
  for (int c = 0; c < n_prt; ++c)
    for (int q = 0; q < P; ++q) {
      const double j = c == 0 ? 0.0 : 0.3;
      pr.x[c][q] = (float)(pts[q].x + j * rng.n());
      pr.y[c][q] = (float)(pts[q].y + j * rng.n());
      pr.z[c][q] = (float)(pts[q].z + j * rng.n());
    }
  finish_protein(*h, pr, *ene);
  // per-point charges ~N(0, 0.3), net zero
  {
    double s = 0;
    for (int q = 0; q < P; ++q) {
      h->prt_ele[q] = (float)(0.3 * rng.n());
      s += h->prt_ele[q];
    }
    for (int q = 0; q < P; ++q) h->prt_ele[q] = (float)(h->prt_ele[q] - s / P);
  }
  }  // protein draws

  // ---- KDE: 120 per atom
  std::vector<int> kt;
  std::vector<float> kx, ky, kz;
  for (int n = 0; n < 120 * L; ++n) {
    const int j = rng.i(0, L - 1);
    kt.push_back(lt[j]);
    kx.push_back((float)(native[j].x + rng.n()));
    ky.push_back((float)(native[j].y + rng.n()));
    kz.push_back((float)(native[j].z + rng.n()));
  }
  finish_kde(*h, kt, kx, ky, kz);
  // lt is in generation (file) order; kde types match the ligand's

  // ---- MCS: 64 rows over the type-sorted atom order
  h->M = 64;
  h->mcs_x.assign((size_t)64 * L, GD_MCS_INVALID);
  h->mcs_y = h->mcs_x;
  h->mcs_z = h->mcs_x;
  h->mcs_tcc.resize(64);
  h->mcs_ncol.resize(64);
  for (int m = 0; m < 64; ++m) {
    int nmatch = 0;
    for (int j = 0; j < L; ++j) {
      if (rng.u() >= 0.5 && !(j == L - 1 && nmatch == 0)) continue;
      const int f = h->lig_perm[j];
      h->mcs_x[(size_t)m * L + j] = (float)(native[f].x + rng.n());
      h->mcs_y[(size_t)m * L + j] = (float)(native[f].y + rng.n());
      h->mcs_z[(size_t)m * L + j] = (float)(native[f].z + rng.n());
      ++nmatch;
    }
    h->mcs_tcc[m] = (float)(0.1 + 0.9 * rng.u());
    h->mcs_ncol[m] = nmatch;
  }

  // ---- PSP: 32 entries on the residues nearest the pocket
  std::map<std::pair<int, int>, float> psp;
  {
    std::vector<std::pair<double, int>> dist;
    for (int i = 0; i < pr.pnr; ++i) dist.push_back({norm(cas[i] - center), pr.seq3[i]});
    std::sort(dist.begin(), dist.end());
    for (int k = 0; k < 32 && k < (int)dist.size(); ++k)
      psp[{dist[k].second, lt[rng.i(0, L - 1)]}] = (float)rng.n();
  }
  set_psp(*h, psp);

  h->n_tmp = n_tmp;
  h->minus_beta.resize(n_tmp);
  set_temperature(n_tmp, floor_temp, ceiling_temp, h->minus_beta.data());
  h->raw_conf = n_lig - 1;
  *out = h;
  return GD_OK;
}

extern "C" gd_status gd_synth_complex(int P, int L, int n_lig, int n_prt, int n_tmp,
                                      float floor_temp, float ceiling_temp, uint64_t seed,
                                      const gd_enepara *ene, gd_host_job **out) {
  Rng rng(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull);
  return synth(P, L, n_lig, n_prt, n_tmp, floor_temp, ceiling_temp, rng, rng, 0.0, ene, out);
}

// A complex of a virtual screen: the protein of protein_seed (the same target
// for every ligand_seed, its pocket a fixed 14 A cavity: the generator's
// 80-atom ligands reach about 11.5 A from their centroid) with the ligand of
// ligand_seed.
extern "C" gd_status gd_synth_screen(int P, uint64_t protein_seed, int L, uint64_t ligand_seed,
                                     int n_lig, int n_prt, int n_tmp, float floor_temp,
                                     float ceiling_temp, const gd_enepara *ene, gd_host_job **out) {
  Rng rl(ligand_seed * 0x9E3779B97F4A7C15ull + 0x1234567ull);
  Rng rp(protein_seed * 0xD1B54A32D192ED03ull + 0x7F4A7C15ull);
  return synth(P, L, n_lig, n_prt, n_tmp, floor_temp, ceiling_temp, rl, rp, 14.0, ene, out);
}

extern "C" gd_status gd_host_job_desc(const gd_host_job *h, gd_job_desc *d) {
  if (!h || !d) return GD_ERR_ARG;
  std::memset(d, 0, sizeof(*d));
  d->n_lig = h->n_lig;
  d->n_prt = h->n_prt;
  d->n_tmp = h->n_tmp;
  d->lig_natom = h->L;
  d->prt_npoint = h->P;
  d->kde_npoint = h->K;
  d->mcs_nrow = h->M;
  d->lig_x = h->lig_x.data();
  d->lig_y = h->lig_y.data();
  d->lig_z = h->lig_z.data();
  d->lig_t = h->lig_t.data();
  d->lig_c = h->lig_c.data();
  d->lig_perm = h->lig_perm.data();
  for (int i = 0; i < 3; ++i) {
    d->lig_center[i] = h->lig_center[i];
    d->pocket_center[i] = h->pocket[i];
  }
  d->prt_x = h->prt_x.data();
  d->prt_y = h->prt_y.data();
  d->prt_z = h->prt_z.data();
  d->prt_t = h->prt_t.data();
  d->prt_c = h->prt_c.data();
  d->prt_cdc = h->prt_cdc.data();
  d->prt_seq3r = h->prt_seq3r.data();
  d->prt_hpp = h->prt_hpp.data();
  d->prt_ele = h->prt_ele.data();
  d->psp = h->psp.data();
  d->psp_ncol = h->psp_ncol;
  d->kde_t = h->kde_t.data();
  d->kde_x = h->kde_x.data();
  d->kde_y = h->kde_y.data();
  d->kde_z = h->kde_z.data();
  d->mcs_x = h->mcs_x.data();
  d->mcs_y = h->mcs_y.data();
  d->mcs_z = h->mcs_z.data();
  d->mcs_tcc = h->mcs_tcc.data();
  d->ene = &h->ene;
  d->minus_beta = h->minus_beta.data();
  return GD_OK;
}

extern "C" gd_status gd_host_job_free(gd_host_job *h) {
  delete h;
  return GD_OK;
}

extern "C" float gd_kabsch_rmsd(int n, const float *a, const float *b) { return kabsch_rmsd(n, a, b); }

extern "C" gd_status gd_set_temperature(int n, float floor_t, float ceil_t, float *mb) {
  if (n < 1 || !mb) return GD_ERR_ARG;
  set_temperature(n, floor_t, ceil_t, mb);
  return GD_OK;
}

// Loader diagnostics for tests: raw ensemble size, per-conformer RMSD to the
// native pose, declared matches per MCS row.
extern "C" gd_status gd_host_job_diag(const gd_host_job *h, int *raw_conf, float *rmsd,
                                      int rmsd_cap, int *mcs_ncol, int ncol_cap) {
  if (!h) return GD_ERR_ARG;
  if (raw_conf) *raw_conf = h->raw_conf;
  if (rmsd)
    for (int i = 0; i < rmsd_cap && i < (int)h->conf_rmsd.size(); ++i) rmsd[i] = h->conf_rmsd[i];
  if (mcs_ncol)
    for (int i = 0; i < ncol_cap && i < (int)h->mcs_ncol.size(); ++i) mcs_ncol[i] = h->mcs_ncol[i];
  return GD_OK;
}
