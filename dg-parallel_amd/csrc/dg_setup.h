// Host-side FP64 construction of the element tables for the triangle kernels
// (what DGAnalysis's constructor builds per cell and face, DGAnalysis.cc:133-302,
// DGCell.cc:127-182, 423-555), independent of any device.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dg_capi.h"
#include "dg_laws.h"

namespace dgp {

struct TriSetup {
  int E = 0, P = 0, NP = 0, NF = 0, law = 0;
  LawParams lp{};
  std::vector<int32_t> perm;   // internal index -> caller element
  std::vector<int32_t> faces;  // (eL, lL, eR, lR) in caller numbering
  std::vector<double> xyz;     // caller corner coordinates [E][3][2]
  // structure-of-arrays element table in internal order
  std::vector<double> mg, detj, size;
  std::vector<int32_t> link;
  // per-order tables
  std::vector<double> tab, vtx;
  std::vector<float> tabf;
};

dg_status build_tri_setup(const dg_mesh_desc* m, const dg_law_desc* law, const dg_options* opt, TriSetup& S,
                          std::string& err);

}  // namespace dgp
