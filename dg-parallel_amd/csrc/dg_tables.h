// Discretisation tables built on the host in FP64: orthonormal modal bases on
// the reference simplices, the quadrature rules the reference selects, and the
// per-order blocks the kernels read (basis / gradient values at volume points,
// trace values at face points).
#pragma once
#include <vector>

namespace dgp {

constexpr int MAXP_TRI = 4;
constexpr int MAXP_TET = 3;

inline constexpr int np_tri(int p) { return (p + 1) * (p + 2) / 2; }
inline constexpr int np_tet(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
// volume rule sizes for degree 2p (GaussQuadratureTri.cc:140, GaussQuadratureTet.cc:351)
inline constexpr int nqv_tri(int p) { return p == 0 ? 1 : p == 1 ? 3 : p == 2 ? 6 : p == 3 ? 12 : 16; }
inline constexpr int nqv_tet(int p) { return p == 0 ? 1 : p == 1 ? 4 : p == 2 ? 14 : 24; }
// face rule size for face degree 2p+1: Gauss-Legendre p+1 points on edges
// (getNGQ1DPts, GaussLegendre1D.cc:271-274), triangle rule GQT[2p+1] on tet faces
inline constexpr int nqf_tri(int p) { return p + 1; }
inline constexpr int nqf_tet(int p) { return p == 0 ? 1 : p == 1 ? 4 : p == 2 ? 7 : 13; }

struct QPoint {
  double x[3];
  double w;  // scaled by the reference measure (1/2 triangle, 1/6 tet, edges: sum 2)
};

std::vector<QPoint> tri_rule(int degree);
std::vector<QPoint> tet_rule(int degree);
std::vector<QPoint> gl_rule(int npts);

// reference coordinates of the point s in [-1, 1] of local edge l (v_l, v_{l+1}):
// (1+s)/2 v_l + (1-s)/2 v_{l+1}, the face points of the trace tables
void edge_point(int l, double s, double* xi);

// phi[k], dphi[k*dim + d] of the orthonormal basis of order p
void basis(int dim, int p, const double* xi, double* phi, double* dphi);

// ---------------------------------------------------------------------------
// Per-order table block read by the triangle kernels (all FP64, converted to
// FP32 for the reduced form).  Layout for order P with NQ = nqv_tri(P),
// NP = np_tri(P), NF = nqf_tri(P):
//   phi  [NQ][NP]          basis at volume points
//   dphi [NQ][NP][2]       reference gradients at volume points
//   wv   [NQ]              volume weights (x 1/2)
//   phif [3][NF][NP]       traces on local edge l at Gauss-Legendre point q,
//                          point x(s) = (1+s)/2 v_l + (1-s)/2 v_{l+1}
//   wf   [NF]              Gauss-Legendre weights x 1/2 (edge Jacobian L/2)
// Mixed-order neighbours use the face rule of the larger order, so each block
// also holds traces for every face rule size nq in [P+1, MAXP_TRI+1]:
//   phifx, for nq = P+2 .. MAXP_TRI+1 in turn: [3][nq][NP]
struct TriBlock {
  static constexpr int off_phi(int) { return 0; }
  static constexpr int off_dphi(int P) { return nqv_tri(P) * np_tri(P); }
  static constexpr int off_wv(int P) { return off_dphi(P) + 2 * nqv_tri(P) * np_tri(P); }
  static constexpr int off_phif(int P) { return off_wv(P) + nqv_tri(P); }
  static constexpr int off_wf(int P) { return off_phif(P) + 3 * nqf_tri(P) * np_tri(P); }
  static constexpr int off_phifx(int P, int nq) {
    int o = off_wf(P) + nqf_tri(P);
    for (int m = P + 2; m < nq; m++) o += 3 * m * np_tri(P);
    return o;
  }
  static constexpr int size(int P) { return off_phifx(P, MAXP_TRI + 2); }
};
// Gauss-Legendre weights (x 1/2) for every rule size used by faces
constexpr int GLW_OFF(int nq) { return (nq - 1) * nq / 2; }

// The trace tables of every order pn (1..4) and face rule size nq (pn+1 ..
// MAXP_TRI+1), [3][nq][Np(pn)] each (the phif / phifx rows of the order-pn
// block), concatenated pn-major: the mixed-order kernel stages them in LDS and
// reads a neighbour's table per lane.
constexpr int ntab_off(int pn, int nq) {
  const int base = pn == 1 ? 0 : pn == 2 ? 126 : pn == 3 ? 342 : 612;  // the blocks of lower orders
  return base + 3 * ((pn + 1) * (pn + 2) / 2) * ((pn + nq) * (nq - 1 - pn) / 2);
}
constexpr int MIX_NTAB = 837;
constexpr bool ntab_layout_ok() {
  int o = 0;
  for (int p = 1; p <= MAXP_TRI; p++)
    for (int m = p + 1; m <= MAXP_TRI + 1; m++) {
      if (ntab_off(p, m) != o) return false;
      o += 3 * m * np_tri(p);
    }
  return o == MIX_NTAB;
}
static_assert(ntab_layout_ok(), "trace table layout");

// ---------------------------------------------------------------------------
// Per-order block of the tetrahedron kernels (Maxwell: linear flux, so the
// volume term is the exact operator form of the degree-2P quadrature).  With
// NP = np_tet(P), NQF = nqf_tet(P) (triangle rule of degree 2P+1 on each face):
//   K    [3][NP][NP]       K_r[k][j] = int d phi_k/d xi_r phi_j over the reference tet
//   phif [4][NQF][NP]      traces on local face l at face point q; the point is
//                          sum_i lambda_i(q) v_{f(l,i)} with f(l,.) the face's local
//                          vertices and lambda = (1-u-v, u, v) of the rule point (u,v)
//   wf   [NQF]             face weights, summing to 1 (physical area x wf)
//   phin [4][6][NQF][NP]   the neighbour's traces at the same points: its local
//                          face l' whose vertex f(l',perm_o(i)) holds the own
//                          face's vertex i, for the six permutations perm_o
struct TetBlock {
  static constexpr int off_K(int) { return 0; }
  static constexpr int off_phif(int P) { return 3 * np_tet(P) * np_tet(P); }
  static constexpr int off_wf(int P) { return off_phif(P) + 4 * nqf_tet(P) * np_tet(P); }
  static constexpr int off_phin(int P) { return off_wf(P) + nqf_tet(P); }
  static constexpr int size(int P) { return off_phin(P) + 4 * 6 * nqf_tet(P) * np_tet(P); }
};
// local faces of a tet (vertex triples) and the six permutations of a face's vertices
extern const int TET_FACE_V[4][3];
extern const int PERM3[6][3];
void build_tet_block(int P, double* out);

// fill one block (doubles); out must hold TriBlock::size(P)
void build_tri_block(int P, double* out);
// Gauss-Legendre half-weights for nq = 1..MAXP_TRI+1, concatenated
void build_glw(double* out);

}  // namespace dgp
