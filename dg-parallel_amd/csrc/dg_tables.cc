// Host FP64 discretisation tables.  The rule and Gram-Schmidt data come from
// dg_rules_data.h (generated by gen_rules.py); the selection of rules by degree
// follows Integrator.cc:31-101 and the basis evaluation follows
// ShapeFunctionList.cc:120-175 (phi_i = sum_j C[i][j] * monomial_j).
#include "dg_tables.h"

#include <cstring>
#include <stdexcept>

#include "dg_rules_data.h"

namespace dgp {

std::vector<QPoint> tri_rule(int degree) {
  if (degree < 0 || degree > DG_TRI_RULE_MAXDEG) throw std::runtime_error("triangle rule degree out of range");
  std::vector<QPoint> r;
  int o = dg_tri_rule_off[degree], n = dg_tri_rule_npt[degree];
  for (int q = 0; q < n; q++)
    r.push_back({{dg_tri_rule_pts[o + q][0], dg_tri_rule_pts[o + q][1], 0.0}, 0.5 * dg_tri_rule_pts[o + q][2]});
  return r;
}

std::vector<QPoint> tet_rule(int degree) {
  if (degree < 0 || degree > DG_TET_RULE_MAXDEG) throw std::runtime_error("tet rule degree out of range");
  std::vector<QPoint> r;
  int o = dg_tet_rule_off[degree], n = dg_tet_rule_npt[degree];
  for (int q = 0; q < n; q++)
    r.push_back({{dg_tet_rule_pts[o + q][0], dg_tet_rule_pts[o + q][1], dg_tet_rule_pts[o + q][2]},
                 dg_tet_rule_pts[o + q][3] / 6.0});
  return r;
}

std::vector<QPoint> gl_rule(int npts) {
  if (npts < 1 || npts > DG_GL_MAXPTS) throw std::runtime_error("Gauss-Legendre size out of range");
  std::vector<QPoint> r;
  int o = dg_gl_off[npts - 1];
  for (int q = 0; q < npts; q++) r.push_back({{dg_gl_pts[o + q][0], 0.0, 0.0}, dg_gl_pts[o + q][1]});
  return r;
}

static double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

void basis(int dim, int p, const double* xi, double* phi, double* dphi) {
  int n = dim == 2 ? np_tri(p) : np_tet(p);
  double m[DG_TET_MAXMODES], dm[DG_TET_MAXMODES][3];
  for (int j = 0; j < n; j++) {
    if (dim == 2) {
      int a = dg_tri_mono[j][0], b = dg_tri_mono[j][1];
      m[j] = ipow(xi[0], a) * ipow(xi[1], b);
      dm[j][0] = a ? a * ipow(xi[0], a - 1) * ipow(xi[1], b) : 0.0;
      dm[j][1] = b ? b * ipow(xi[0], a) * ipow(xi[1], b - 1) : 0.0;
      dm[j][2] = 0.0;
    } else {
      int a = dg_tet_mono[j][0], b = dg_tet_mono[j][1], c = dg_tet_mono[j][2];
      m[j] = ipow(xi[0], a) * ipow(xi[1], b) * ipow(xi[2], c);
      dm[j][0] = a ? a * ipow(xi[0], a - 1) * ipow(xi[1], b) * ipow(xi[2], c) : 0.0;
      dm[j][1] = b ? b * ipow(xi[0], a) * ipow(xi[1], b - 1) * ipow(xi[2], c) : 0.0;
      dm[j][2] = c ? c * ipow(xi[0], a) * ipow(xi[1], b) * ipow(xi[2], c - 1) : 0.0;
    }
  }
  for (int i = 0; i < n; i++) {
    const double* C = dim == 2 ? dg_tri_gs[i] : dg_tet_gs[i];
    double r = 0, g0 = 0, g1 = 0, g2 = 0;
    for (int j = 0; j <= i; j++) {
      r += C[j] * m[j];
      g0 += C[j] * dm[j][0];
      g1 += C[j] * dm[j][1];
      g2 += C[j] * dm[j][2];
    }
    if (phi) phi[i] = r;
    if (dphi) {
      dphi[i * dim + 0] = g0;
      dphi[i * dim + 1] = g1;
      if (dim == 3) dphi[i * dim + 2] = g2;
    }
  }
}

void edge_point(int l, double s, double* xi) {
  static const double V[3][2] = {{0, 0}, {1, 0}, {0, 1}};
  double a = 0.5 * (1.0 + s), b = 0.5 * (1.0 - s);
  const double* A = V[l];
  const double* B = V[(l + 1) % 3];
  xi[0] = a * A[0] + b * B[0];
  xi[1] = a * A[1] + b * B[1];
  xi[2] = 0.0;
}

void build_tri_block(int P, double* out) {
  const int NP = np_tri(P), NQ = nqv_tri(P), NQF = nqf_tri(P);
  std::memset(out, 0, sizeof(double) * TriBlock::size(P));
  auto vr = tri_rule(2 * P);
  if ((int)vr.size() != NQ) throw std::runtime_error("volume rule size mismatch");
  double phi[64], dphi[128];
  for (int q = 0; q < NQ; q++) {
    basis(2, P, vr[q].x, phi, dphi);
    for (int k = 0; k < NP; k++) {
      out[TriBlock::off_phi(P) + q * NP + k] = phi[k];
      out[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 0] = dphi[k * 2 + 0];
      out[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 1] = dphi[k * 2 + 1];
    }
    out[TriBlock::off_wv(P) + q] = vr[q].w;
  }
  for (int nq = NQF; nq <= MAXP_TRI + 1; nq++) {
    auto gl = gl_rule(nq);
    int base = nq == NQF ? TriBlock::off_phif(P) : TriBlock::off_phifx(P, nq);
    for (int l = 0; l < 3; l++)
      for (int q = 0; q < nq; q++) {
        double xi[3];
        edge_point(l, gl[q].x[0], xi);
        basis(2, P, xi, phi, nullptr);
        for (int k = 0; k < NP; k++) out[base + (l * nq + q) * NP + k] = phi[k];
      }
    if (nq == NQF)
      for (int q = 0; q < nq; q++) out[TriBlock::off_wf(P) + q] = 0.5 * gl[q].w;
  }
}

void build_glw(double* out) {
  for (int nq = 1; nq <= MAXP_TRI + 1; nq++) {
    auto gl = gl_rule(nq);
    for (int q = 0; q < nq; q++) out[GLW_OFF(nq) + q] = 0.5 * gl[q].w;
  }
}

// local face l = vertices {0,1,2}, {0,1,3}, {1,2,3}, {0,2,3} (the oracle's and the
// mesh generator's convention; face l is opposite vertex 3, 2, 0, 1)
const int TET_FACE_V[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};
const int PERM3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

void build_tet_block(int P, double* out) {
  const int NP = np_tet(P), NQF = nqf_tet(P);
  std::memset(out, 0, sizeof(double) * TetBlock::size(P));
  static const double V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double phi[64], dphi[192];
  // stiffness operators with the rule of degree 2P (integrand degree 2P-1)
  auto vr = tet_rule(2 * P);
  for (const auto& q : vr) {
    basis(3, P, q.x, phi, dphi);
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < NP; k++)
        for (int j = 0; j < NP; j++) out[TetBlock::off_K(P) + (r * NP + k) * NP + j] += q.w * dphi[k * 3 + r] * phi[j];
  }
  auto fr = tri_rule(2 * P + 1);
  if ((int)fr.size() != NQF) throw std::runtime_error("tet face rule size mismatch");
  for (int q = 0; q < NQF; q++) out[TetBlock::off_wf(P) + q] = 2.0 * fr[q].w;
  for (int l = 0; l < 4; l++)
    for (int q = 0; q < NQF; q++) {
      const double lam[3] = {1.0 - fr[q].x[0] - fr[q].x[1], fr[q].x[0], fr[q].x[1]};
      double xi[3] = {0, 0, 0};
      for (int i = 0; i < 3; i++)
        for (int d = 0; d < 3; d++) xi[d] += lam[i] * V[TET_FACE_V[l][i]][d];
      basis(3, P, xi, phi, nullptr);
      for (int k = 0; k < NP; k++) out[TetBlock::off_phif(P) + (l * NQF + q) * NP + k] = phi[k];
      for (int o = 0; o < 6; o++) {
        double xn[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++)
          for (int d = 0; d < 3; d++) xn[d] += lam[i] * V[TET_FACE_V[l][PERM3[o][i]]][d];
        basis(3, P, xn, phi, nullptr);
        for (int k = 0; k < NP; k++) out[TetBlock::off_phin(P) + ((l * 6 + o) * NQF + q) * NP + k] = phi[k];
      }
    }
}

}  // namespace dgp
