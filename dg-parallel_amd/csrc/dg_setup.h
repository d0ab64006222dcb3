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

// What order changes keep: face connectivity, boundary codes and geometry of
// every element, in the caller's numbering.
// A face of half an edge on a mesh with hanging nodes: the child's local face
// lc lies on half h of the coarse cell K's local face lK (half 0 holds K's
// vertex lK); rev: the child runs along the face against K's direction.
struct HangFace {
  int child, lc, K, lK, h, rev;
};

struct TriTopo {
  int E = 0;
  std::vector<HangFace> hang;           // faces of half an edge (1-irregular meshes)
  std::vector<int32_t> nbe, nbl, same;  // [E*3]: neighbour, its local face, equal first vertex (-1: none)
  std::vector<int8_t> bcode;            // [E*3]: BC_* of a face without neighbour
  std::vector<int32_t> sig;             // [E]: link signature (faces' neighbour face, orientation, boundary)
  std::vector<double> mg, detj, size;   // [E*4] |J| J^-1 (r*2+d), [E], [E] (2 x inradius)
  std::vector<int32_t> faces;           // (eL, lL, eR, lR)
};

// a triangle's size, 2 x inradius (DGCell::getSize, DGCell.cc:1209-1229), from its
// corners [3][2]
double tri_cell_size(const double* X);

// Caller topology and geometry (mMesh.cc:321-378, Mapping.cc:442-566,
// DGCell.cc:1209-1229); validates the mesh.  A face (a, b) without a
// neighbour, where a vertex m closes two other such faces (a, m) and (m, b)
// with m - a = b - m (each in its own element's coordinates), is a coarse
// cell's face split on the other side (red refinement, mDGMesh.cc:54-337): the
// two halves become faces of half an edge (hang, faces list: child left, coarse
// cell right), and the three faces get the code BC_HANG.
dg_status build_tri_topo(const dg_mesh_desc* m, TriTopo& T, std::string& err);

struct TriSetup {
  int E = 0, P = 0, NP = 0, NF = 0, law = 0;
  TriTopo topo;
  // Elements are bucketed by order: internal elements [bstart[p], bstart[p+1])
  // have order p (p = 1..4) and their coefficients start at double boff[p] of
  // the state arrays, stored [row][element] within the bucket.  A uniform-order
  // mesh is one bucket at offset 0 (P, NP describe it); mixed is set otherwise.
  bool mixed = false;
  int bstart[6] = {0, 0, 0, 0, 0, 0};
  long boff[6] = {0, 0, 0, 0, 0, 0};
  std::vector<int8_t> order;    // internal order of each element
  std::vector<long> coff;       // caller element -> first coefficient in the caller layout (E+1)
  // mixed-order tables: the blocks of orders 1..4 at moff[p]; ftoff[p][nq] is the
  // trace table [3][nq][Np(p)] of order p for the face rule of nq points;
  // glw the Gauss-Legendre half weights (GLW_OFF); mvtx[p][4][15] the mean and
  // vertex basis values of order p
  std::vector<double> mtab, glw, mvtx;
  std::vector<double> ntab;  // every order's trace tables, MIX_NTAB doubles at ntab_off(p, nq)
  int moff[6] = {0, 0, 0, 0, 0, 0};
  int ftoff[6][6];
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
                          std::string& err, const int32_t* grp = nullptr);

// The internal numbering and element tables of S for per-element orders ord
// (caller numbering, 1..4) over topology T: perm, order, bucket bounds, coff,
// mg, detj, size, link and the mixed flag (S.E, S.NF set by the caller).
// Internal order: order buckets, then consecutive blocks of 128 caller elements,
// then link signature (stable), so a wavefront takes one path and face
// neighbours sit in nearby tiles.  grp (caller numbering, uniform orders only):
// a group key ahead of the blocks (the classes of local time stepping), so each
// group is one internal range.  geom = false leaves S.mg, S.detj and S.size
// unset (a caller that moves the geometry itself, as re-bucketing on the device).
void tri_layout(const TriTopo& T, const int* ord, TriSetup& S, const int32_t* grp = nullptr, bool geom = true);

// The hanging-face pass's tables for S (one order): the owning elements of faces
// of half an edge (internal numbering), records [4] per face and side (own
// local face, the other side, its local face, flags: bit 0 own side coarse,
// bit 1 reversed points, bit 2 the half) grouped by element (rstart), and the
// order-P basis at the Gauss points of each face's halves [3][2][NQF][NP]
// (each half's points in the face's direction).
void build_hang_tables(const TriSetup& S, std::vector<int32_t>& elem, std::vector<int32_t>& rstart,
                       std::vector<int32_t>& rec, std::vector<double>& ph);

// Tetrahedra (Maxwell, one order): the same roles as TriSetup.  mg holds
// |J| J^-1 as [9][E] (row-major r*3+d), size the shortest edge (DGCell.cc:1230-1262),
// link[4][E] = (nb << 5) | (perm << 2) | nb_local_face with perm indexing
// PERM3 (the neighbour's face vertex PERM3[perm][i] holds the own face vertex i),
// or -(boundary code).
struct TetSetup {
  int E = 0, P = 0, NP = 0, NF = 0, law = 0;
  LawParams lp{};
  std::vector<int32_t> perm, faces;
  std::vector<double> xyz;  // [E][4][3]
  std::vector<double> mg, detj, size;
  std::vector<int32_t> link;
  std::vector<double> tab;
};

// a field (dg_field_fn, time 0) at every volume point and every face point of
// every element: [NQV + 3 NQF][nval][E] in the internal order (element i is
// caller element perm[i], corners xyz [E][3][2] in caller order), face l point q
// at row NQV + l NQF + q (the points of the order-P trace tables); false when a
// value is not finite
bool tri_point_field(const std::vector<double>& xyz, const std::vector<int32_t>& perm, int P, int nval,
                     dg_field_fn fn, void* user, std::vector<double>& out);

dg_status build_tet_setup(const dg_mesh_desc* m, const dg_law_desc* law, const dg_options* opt, TetSetup& S,
                          std::string& err);

}  // namespace dgp
