// Mixed-order (P1-P4) triangle kernels as per-element __host__ __device__
// functions, shared by the HIP kernels (dg_mixed.hip) and the CPU emulation (the
// mixed-order stage itself is the tiled kernel of dg_mixed_tiled.h):
//
//   limit_element        triangle TVB limiter (new design; the reference's
//                        DGMomentLimiter::limit, DGLimiter.cc:963-1133, indexes
//                        quadrilateral tensor modes only);
//   sensor_element       modal-decay shock sensor choosing the next order (new
//                        design in place of the face-jump bins of DGPref.cc:24-60,
//                        194-228, 602-630);
//   dt_element           dt_e = CFL size_e / ((2p_e+1) lambda_e) (DGCell.cc:1155-1262).
//
// State layout: bucket p (order p) holds internal elements [bstart[p], bstart[p+1])
// and its coefficients from double boff[p] on, [row = f*Np(p)+k][element].
#pragma once

#include <cmath>
#include <cstdint>

#include "dg_kernels.h"
#include "dg_laws.h"
#include "dg_stage_tiled.h"
#include "dg_tables.h"

namespace dgp {

constexpr int MIX_CHUNK = 64;   // elements per work item of the persistent launch (one tile)
constexpr int MIX_MAXNP = 15;   // np_tri(4)

struct MixedArgs {
  int E;
  int bstart[6];
  long boff[6];
  const double* tab;  // blocks of orders 1..4
  int moff[6];
  int ftoff[6][6];    // [p][nq]: trace table [3][nq][Np(p)] (absolute offset into tab)
  const double* glw;  // Gauss-Legendre half weights at GLW_OFF(nq)
  const double* mvtx; // [p][4][15]: phi_0 at [p][0][0], then the 3 vertices
  TriElemTable el;
  LawParams law;
  const double* Uin;
  const double* Un;
  double* acc;
  double* Uout;
  double* R;
  double dt;
  int scheme, stage;  // scheme MODE_RESIDUAL: write R only
  double* resid_part; // per workgroup of the persistent launch (or null)
  const int* chunks;  // work items: (order, first internal element) pairs
  int nchunks;
  const double* ntab; // neighbour trace tables of every order and rule size (dg_mixed_tiled.h)
};

// order and coefficient index from the bucket bounds of any argument block
// (MixedArgs, LimiterArgs, SensorArgs all carry bstart/boff)
template <class A>
DG_HD int mixed_order_of(const A& a, int i) {
  return 1 + (i >= a.bstart[2]) + (i >= a.bstart[3]) + (i >= a.bstart[4]);
}
template <class A>
DG_HD long mixed_index(const A& a, int p, int i, int row) {
  return a.boff[p] + (long)row * (a.bstart[p + 1] - a.bstart[p]) + (i - a.bstart[p]);
}

// outward unit normal and length of local edge l from |J| J^-1 (as the uniform kernels)
DG_HD void edge_normal(const double mg[2][2], int l, double n[2], double& len) {
  double nl0, nl1;
  if (l == 0) {
    nl0 = -mg[1][0];
    nl1 = -mg[1][1];
  } else if (l == 1) {
    nl0 = mg[0][0] + mg[1][0];
    nl1 = mg[0][1] + mg[1][1];
  } else {
    nl0 = -mg[0][0];
    nl1 = -mg[0][1];
  }
  len = dg_sqrt(nl0 * nl0 + nl1 * nl1);
  n[0] = nl0 / len;
  n[1] = nl1 / len;
}

// ---------------------------------------------------------------------------
// Triangle TVB limiter (new design), per element, Jacobi semantics: it reads the
// means (mode 0, never modified) of its neighbours and rewrites only its own
// modes 1.. in place, so no ordering between elements matters.  Euler elements
// then pass the positivity scaling below (flag bit 4).
//
// For field f, with the element mean m = U_0 phi_0 and the face-neighbour means
// m_l (boundary faces: the law's ghost state of the mean), at the midpoint c_l
// of edge l:
//   u_l   = u(c_l) - m                          the whole polynomial's deviation
//   d_l   = U_1 phi_1(c_l) + U_2 phi_2(c_l)    the linear part's deviation
//   D_l   = m_l - m
//   the field is troubled when some |u_l| > tol and the minmod moves u_l:
//   |u_l - minmod(u_l, nu D_l)| > 1e-10 (|u_l| + |nu D_l|) (a relative margin:
//   a slope the limiter rebuilt equals the neighbour difference exactly, a tie
//   that the next stage keeps to rounding); tol = max(M h^2, 1e-10 s) with the
//   field's scale s = |m|, sqrt(|rho E|) for the Euler momenta (h = 2 x
//   inradius; M = 0 is the TVD limiter; Cockburn & Shu 1998, sec. 5.3, for P^k);
// a troubled field keeps its mean, takes the linear part rebuilt from the
// limited deviations dl_l = minmod(d_l, nu D_l) (after the Cockburn-Shu
// positive/negative rescaling that makes them sum to zero) and drops every
// mode of degree >= 2.  Bit f of the returned flag is set when field f was limited.
struct LimiterArgs {
  int E;
  int bstart[6];
  long boff[6];
  TriElemTable el;
  LawParams law;
  double* U;          // limited in place
  const double* mvtx; // phi_0 at [p*60]
  double midv[3][MIX_MAXNP];  // the basis (hierarchical: every order's prefix) at the three edge midpoints
  double pinv[2][3];  // linear coefficients from midpoint deviations
  double tvb_m, nu;
  uint8_t* flags;     // internal order
  // positivity check points (Euler): the blocks of orders 1..4 (volume points,
  // edge points of every face rule an element of that order meets)
  const double* tab;
  int moff[6];
  int ftoff[6][6];
};

DG_HD double lim_minmod(double a, double b) {
  if (a > 0.0 && b > 0.0) return a < b ? a : b;
  if (a < 0.0 && b < 0.0) return a > b ? a : b;
  return 0.0;
}

// Positivity of density and pressure (Zhang & Shu 2010, J. Comput. Phys. 229,
// 8918-8934), new in this build (the reference has none): after the TVB step,
// the modes above the mean of an Euler element are scaled towards the mean so
// that density and then pressure stay >= eps = min(1e-13, mean density, mean
// pressure) at every point the scheme evaluates: the volume points of its order,
// the edge points of every face rule it meets and its vertices.  Returns true
// when it scaled.
DG_HD double pos_pressure(double gm1, const double* w) {
  return gm1 * (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
}

// one check point (basis row without phi_0) of the positivity pass; pass 0
// tracks the density bound, pass 1 the pressure bound with density scaled by th1
template <int NP>
DG_HD void pos_point(const double* row, const double (*U)[NP], const double* mean, int pass, double gm1, double eps,
                     double& th1, double& th2) {
  if (pass == 0) {
    double acc = 0.0;
#pragma unroll
    for (int k = 1; k < NP; k++) acc += row[k] * U[0][k];
    const double w0 = mean[0] + acc;
    if (w0 < eps) {
      const double t = (mean[0] - eps) / (mean[0] - w0);
      if (t < th1) th1 = t;
    }
    return;
  }
  double w[4];
#pragma unroll
  for (int f = 0; f < 4; f++) {
    double acc = 0.0;
#pragma unroll
    for (int k = 1; k < NP; k++) acc += row[k] * U[f][k];
    w[f] = mean[f] + (f == 0 ? th1 : 1.0) * acc;
  }
  if (pos_pressure(gm1, w) < eps) {
    // p(mean + s (w - mean)) = eps: rho(s) (E(s) - eps/(g-1)) - |m(s)|^2 / 2 = 0
    const double dr = w[0] - mean[0], du = w[1] - mean[1], dv = w[2] - mean[2], de = w[3] - mean[3];
    const double e0 = mean[3] - eps / gm1;
    const double qa = dr * de - 0.5 * (du * du + dv * dv);
    const double qb = dr * e0 + mean[0] * de - (mean[1] * du + mean[2] * dv);
    const double qc = mean[0] * e0 - 0.5 * (mean[1] * mean[1] + mean[2] * mean[2]);
    double t;
    if (qa == 0.0) {
      t = -qc / qb;
    } else {
      const double disc = qb * qb - 4.0 * qa * qc;
      const double sq = dg_sqrt(disc > 0.0 ? disc : 0.0);
      // the root in (0, 1]: qc > 0 at s = 0, f(1) < 0
      const double r1 = (-qb - sq) / (2.0 * qa), r2 = (-qb + sq) / (2.0 * qa);
      t = 1.0;
      if (r1 > 0.0 && r1 < t) t = r1;
      if (r2 > 0.0 && r2 < t) t = r2;
    }
    if (t < 0.0) t = 0.0;
    if (t < th2) th2 = t;
  }
}

// the scaling on the element's coefficients U (registers); true when it scaled.
// Check points in order: the volume points of order P, the vertices, then the
// edge points of the face rules nq = P+1 .. 5 (rows of the order-P trace tables).
template <int P>
DG_HD bool positivity_element(const LimiterArgs& a, double (*U)[np_tri(P)]) {
  constexpr int NP = np_tri(P), NQV = nqv_tri(P);
  const double gm1 = a.law.gamma - 1.0;
  const double phi0 = a.mvtx[P * 60];
  double mean[4];
#pragma unroll
  for (int f = 0; f < 4; f++) mean[f] = U[f][0] * phi0;
  const double pm = pos_pressure(gm1, mean);
  double eps = 1e-13;
  if (mean[0] < eps) eps = mean[0];
  if (pm < eps) eps = pm;
  if (!(eps > 0.0)) return false;  // a non-physical mean: nothing to scale towards
  const double* tb = a.tab + a.moff[P];
  double th1 = 1.0, th2 = 1.0;
  for (int pass = 0; pass < 2; pass++) {
#pragma unroll 1
    for (int q = 0; q < NQV; q++) pos_point<NP>(tb + TriBlock::off_phi(P) + q * NP, U, mean, pass, gm1, eps, th1, th2);
#pragma unroll 1
    for (int v = 0; v < 3; v++) pos_point<NP>(a.mvtx + P * 60 + (1 + v) * 15, U, mean, pass, gm1, eps, th1, th2);
#pragma unroll
    for (int nq = P + 1; nq <= MAXP_TRI + 1; nq++) {
      const double* ft = a.tab + a.ftoff[P][nq];
#pragma unroll 1
      for (int r = 0; r < 3 * nq; r++) pos_point<NP>(ft + r * NP, U, mean, pass, gm1, eps, th1, th2);
    }
  }
  if (th1 >= 1.0 && th2 >= 1.0) return false;
#pragma unroll
  for (int f = 0; f < 4; f++) {
    const double sc = (f == 0 ? th1 : 1.0) * th2;
#pragma unroll
    for (int k = 1; k < NP; k++) U[f][k] = sc * U[f][k];
  }
  return true;
}

// the limiter on one element of order P: its coefficients stay in registers and
// only rewritten fields go back to memory
template <class LAW, int P>
DG_HD void limit_element_p(const LimiterArgs& a, int i) {
  const LimiterArgs& ix = a;
  constexpr int NF = LAW::NF, NP = np_tri(P);
  const int E = a.E;
  const double phi0 = a.mvtx[P * 60];
  double U[NF][NP];
#pragma unroll
  for (int f = 0; f < NF; f++)
#pragma unroll
    for (int k = 0; k < NP; k++) U[f][k] = a.U[mixed_index(ix, P, i, f * NP + k)];
  double mean[NF], full[3][NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    mean[f] = U[f][0] * phi0;
#pragma unroll
    for (int l = 0; l < 3; l++) full[l][f] = U[f][1] * a.midv[l][1];
#pragma unroll
    for (int k = 2; k < NP; k++)
#pragma unroll
      for (int l = 0; l < 3; l++) full[l][f] = fma(U[f][k], a.midv[l][k], full[l][f]);
  }
  double mg[2][2];
  for (int r = 0; r < 2; r++)
    for (int d = 0; d < 2; d++) mg[r][d] = a.el.mg[(long)(r * 2 + d) * E + i];
  double D[3][NF];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    const int link = a.el.link[(long)l * E + i];
    double g[NF];
    if (link >= 0) {
      const int nb = link >> 3, pn = mixed_order_of(ix, nb), npn = (pn + 1) * (pn + 2) / 2;
      for (int f = 0; f < NF; f++) g[f] = a.U[mixed_index(ix, pn, nb, f * npn)] * a.mvtx[pn * 60];
    } else {
      double n[2], len;
      edge_normal(mg, l, n, len);
      LAW::ghost(a.law, -link, n, mean, g);
    }
#pragma unroll
    for (int f = 0; f < NF; f++) D[l][f] = g[f] - mean[f];
  }
  const double h = a.el.size[i];
  unsigned flag = 0;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    double d[3];
    bool troubled = false;
    // the field's scale: its mean, or for the Euler momenta sqrt(rho E) (a
    // momentum that does not vanish in a gas at rest)
    double sc = fabs(mean[f]);
    if constexpr (NF == 4)
      if (f == 1 || f == 2) sc = sqrt(fabs(mean[0] * mean[3]));
    double tol = a.tvb_m * h * h;
    const double tr = 1e-10 * sc;
    if (tr > tol) tol = tr;
#pragma unroll
    for (int l = 0; l < 3; l++) {
      // troubled when the minmod moves the deviation by more than 1e-10 of the
      // compared values: a limited slope equals the neighbour difference exactly,
      // a tie that the next stage keeps to rounding
      const double u = full[l][f], nd = a.nu * D[l][f];
      const double gap = fabs(u - lim_minmod(u, nd));
      if (fabs(u) > tol && gap > 1e-10 * (fabs(u) + fabs(nd))) troubled = true;
      d[l] = fma(U[f][2], a.midv[l][2], U[f][1] * a.midv[l][1]);
    }
    if (!troubled) continue;
    flag |= 1u << f;
    double dl[3], pos = 0.0, neg = 0.0;
#pragma unroll
    for (int l = 0; l < 3; l++) {
      dl[l] = lim_minmod(d[l], a.nu * D[l][f]);
      pos += dl[l] > 0.0 ? dl[l] : 0.0;
      neg += dl[l] < 0.0 ? -dl[l] : 0.0;
    }
    if (pos != neg) {
      const double tp = pos > 0.0 ? (neg < pos ? neg / pos : 1.0) : 0.0;
      const double tn = neg > 0.0 ? (pos < neg ? pos / neg : 1.0) : 0.0;
#pragma unroll
      for (int l = 0; l < 3; l++) dl[l] = dl[l] > 0.0 ? tp * dl[l] : tn * dl[l];
    }
    U[f][1] = fma(a.pinv[0][2], dl[2], fma(a.pinv[0][1], dl[1], a.pinv[0][0] * dl[0]));
    U[f][2] = fma(a.pinv[1][2], dl[2], fma(a.pinv[1][1], dl[1], a.pinv[1][0] * dl[0]));
#pragma unroll
    for (int k = 3; k < NP; k++) U[f][k] = 0.0;
  }
  unsigned dirty = flag;
  if constexpr (NF == 4) {
    if (positivity_element<P>(a, U)) {
      flag |= 16;
      dirty = 15;
    }
  }
#pragma unroll
  for (int f = 0; f < NF; f++)
    if (dirty & (1u << f))
#pragma unroll
      for (int k = 1; k < NP; k++) a.U[mixed_index(ix, P, i, f * NP + k)] = U[f][k];
  a.flags[i] = (uint8_t)flag;
}

template <class LAW>
DG_HD void limit_element(const LimiterArgs& a, int i) {
  switch (mixed_order_of(a, i)) {
    case 1: limit_element_p<LAW, 1>(a, i); break;
    case 2: limit_element_p<LAW, 2>(a, i); break;
    case 3: limit_element_p<LAW, 3>(a, i); break;
    default: limit_element_p<LAW, 4>(a, i); break;
  }
}

// ---------------------------------------------------------------------------
// Modal-decay shock sensor (new design).  On field 0 (density for Euler), with
// the orthonormal modes of the element's order p:
//   top = sum of U_k^2 over the degree-p modes, all = sum over every mode,
//   high = all - U_0^2;
// the next order is p - 1 (not below pmin) when top > s_hi[p] all, p + 1 (not
// above pmax) when top < s_lo[p] all, and p otherwise; an element whose modes
// above the mean carry no energy (high <= 1e-24 all) keeps its order.  Sums run
// in mode order with fused multiply-adds so the oracle can reproduce them bit
// for bit.
struct SensorArgs {
  int E;
  int bstart[6];
  long boff[6];
  const double* U;
  double s_hi[6], s_lo[6];
  int pmin, pmax;
  uint8_t* orders;  // internal order
};

DG_HD void sensor_element(const SensorArgs& a, int i) {
  const SensorArgs& ix = a;
  const int p = mixed_order_of(ix, i), np = (p + 1) * (p + 2) / 2, npl = p * (p + 1) / 2;
  double all = 0.0, top = 0.0, high = 0.0;
  for (int k = 0; k < np; k++) {
    const double u = a.U[mixed_index(ix, p, i, k)];
    all = fma(u, u, all);
    if (k >= 1) high = fma(u, u, high);
    if (k >= npl) top = fma(u, u, top);
  }
  int next = p;
  if (high > 1e-24 * all) {
    if (top > a.s_hi[p] * all)
      next = p - 1 < a.pmin ? a.pmin : p - 1;
    else if (top < a.s_lo[p] * all)
      next = p + 1 > a.pmax ? a.pmax : p + 1;
  }
  a.orders[i] = (uint8_t)next;
}

// ---------------------------------------------------------------------------
// time-step candidate of one element of any order (DGCell.cc:1155-1207)
template <class LAW>
DG_HD double dt_element(const MixedArgs& a, int i, double cfl) {
  constexpr int NF = LAW::NF;
  const int p = mixed_order_of(a, i), np = (p + 1) * (p + 2) / 2;
  const double* vt = a.mvtx + p * 60;
  double mean[NF];
  for (int f = 0; f < NF; f++) mean[f] = a.Uin[mixed_index(a, p, i, f * np)] * vt[0];
  double lam = LAW::maxev(a.law, mean);
  for (int v = 0; v < 3; v++) {
    double val[NF];
    for (int f = 0; f < NF; f++) {
      double s = 0.0;
      for (int k = 0; k < np; k++) s += a.Uin[mixed_index(a, p, i, f * np + k)] * vt[(1 + v) * 15 + k];
      val[f] = s;
    }
    const double b = LAW::maxev(a.law, val);
    lam = b > lam ? b : lam;
  }
  return a.el.size[i] * cfl / (lam * (2. * p + 1.0));
}

// ---------------------------------------------------------------------------
// The state carried over to new order buckets (rebucketing in place): element i
// of the new numbering was element iold[i] of the old one; raised orders get
// zero modes, lowered orders drop their top modes (DGCell::adaptOrder,
// DGCell.cc:1542-1602).  Bucket bounds and offsets of both layouts.
struct RemapArgs {
  int E, NF;
  int bs0[6], bs1[6];
  long bo0[6], bo1[6];
  const int32_t* iold;
  const double* src;
  double* dst;
  // the element geometry carried to the new internal numbering on the device
  // (null: not moved): |J| J^-1 [4][E], |J| [E], size [E]
  const double *mg0, *detj0, *size0;
  double *mg1, *detj1, *size1;
};

// ---------------------------------------------------------------------------
// Non-physical recovery and the state check (DGCell::setToZero, DGCell.cc:251-279,
// triggered from the volume pass, DGCell.cc:198-223, and from the face pass,
// DGCell.cc:639-732 with parallel.cc:57-72).  An Euler element whose state is
// non-physical (Euler2d::isPhysical, EulerLaw.cc:1398-1404: rho < 0 or p < 0)
// at a point the residual evaluates -- its volume points and its own traces at
// the points of each face rule (max(p, p_nb) + 1 Gauss-Legendre points; p + 1
// on boundary faces) -- keeps its means and drops every mode >= 1 (apply), or
// is only reported (check).  Every element is checked on the input state and
// rewrites only its own modes, so the pass is independent of scheduling; the
// reference instead reduces cells one at a time inside its assembly.
// Flag bits OR-ed into flags[i] and *status: NP_POINT a non-physical point
// (reduced when applying), NP_MEAN a non-physical mean (unrecoverable; the
// reference exits, DGCell.cc:218), NP_NONFINITE a NaN or Inf coefficient.
enum { NP_POINT = 1, NP_MEAN = 2, NP_NONFINITE = 4 };

struct RecoverArgs {
  LimiterArgs l;      // the state (in place), tables, links, flags (OR-ed; may be null: report only)
  unsigned* status;   // OR of every element's bits
  int apply;          // 1: reduce; 0: check only
};

DG_HD bool finite_hd(double x) { return x - x == 0.0; }

template <int NF, bool EULER, int P>
DG_HD unsigned recover_element_p(const RecoverArgs& r, int i) {
  const LimiterArgs& a = r.l;
  constexpr int NP = np_tri(P), NQV = nqv_tri(P);
  const int E = a.E;
  double U[NF][NP];
  unsigned bits = 0;
#pragma unroll
  for (int f = 0; f < NF; f++)
#pragma unroll
    for (int k = 0; k < NP; k++) {
      U[f][k] = a.U[mixed_index(a, P, i, f * NP + k)];
      if (!finite_hd(U[f][k])) bits |= NP_NONFINITE;
    }
  if constexpr (EULER) {
    const double gm1 = a.law.gamma - 1.0;
    const double* tb = a.tab + a.moff[P];
    auto bad = [&](const double* row) {
      double w[4];
#pragma unroll
      for (int f = 0; f < 4; f++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < NP; k++) acc += U[f][k] * row[k];
        w[f] = acc;
      }
      return w[0] < 0.0 || gm1 * (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]) < 0.0;
    };
    bool np = false;
#pragma unroll 1
    for (int q = 0; q < NQV && !np; q++) np = bad(tb + TriBlock::off_phi(P) + q * NP);
#pragma unroll 1
    for (int l = 0; l < 3 && !np; l++) {
      const int link = a.el.link[(long)l * E + i];
      int nq = P + 1;
      if (link >= 0) {
        const int pn = mixed_order_of(a, link >> 3);
        if (pn + 1 > nq) nq = pn + 1;
      }
      const double* ft = a.tab + (nq == 2 ? a.ftoff[P][2] : nq == 3 ? a.ftoff[P][3] : nq == 4 ? a.ftoff[P][4]
                                                                                              : a.ftoff[P][5]);
#pragma unroll 1
      for (int q = 0; q < nq && !np; q++) np = bad(ft + (l * nq + q) * NP);
    }
    if (np) {
      bits |= NP_POINT;
      const double phi0 = a.mvtx[P * 60];
      double mean[4];
#pragma unroll
      for (int f = 0; f < 4; f++) mean[f] = U[f][0] * phi0;
      if (mean[0] < 0.0 || gm1 * (mean[3] - 0.5 * (mean[1] * mean[1] + mean[2] * mean[2]) / mean[0]) < 0.0)
        bits |= NP_MEAN;
      if (r.apply)
#pragma unroll
        for (int f = 0; f < 4; f++)
#pragma unroll
          for (int k = 1; k < NP; k++) a.U[mixed_index(a, P, i, f * NP + k)] = 0.0;
    }
  }
  if (bits && a.flags) a.flags[i] = (uint8_t)(a.flags[i] | bits);  // checks without a flag array only report
  return bits;
}

// launches (dg_mixed.hip); part holds one value per workgroup of the launch
int launch_remap(const RemapArgs& a, hipStream_t s);
// the recovery pass or check over every order bucket (triangles; Euler checks
// points, every law checks finiteness)
int launch_recover(const RecoverArgs& a, int law, int nf, hipStream_t s);
int mixed_grid(int nchunks);
int launch_mixed_stage(const MixedArgs& a, int law, int grid, hipStream_t s);
// the bucket-by-bucket form: one launch over bucket p (1..4), one workgroup per
// 64 of its elements, with order p's own registers and LDS; residual partial
// sums at resid_part[part0 ..)
inline int mixed_bucket_tiles(const int* bstart, int p) { return (bstart[p + 1] - bstart[p] + MIX_CHUNK - 1) / MIX_CHUNK; }
int launch_mixed_bucket(const MixedArgs& a, int law, int p, int part0, hipStream_t s);
// part receives one minimum per workgroup: mixed_dt_blocks(E) values
int mixed_dt_blocks(int E);
int launch_mixed_dt(const MixedArgs& a, int law, double cfl, double* part, hipStream_t s);
int launch_limiter(const LimiterArgs& a, int law, hipStream_t s);
// the limiter of one order bucket p; buckets are independent (each rewrites only
// its own elements and reads only neighbour means, which limiting keeps)
int launch_limiter_bucket(const LimiterArgs& a, int law, int p, hipStream_t s);
int launch_sensor(const SensorArgs& a, hipStream_t s);

}  // namespace dgp
