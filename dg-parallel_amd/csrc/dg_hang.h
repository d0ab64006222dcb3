// Faces of half an edge on triangle meshes with hanging nodes (red refinement
// beside unsplit cells, mDGMesh.cc:54-337).  The reference integrates each leaf
// edge with both sides' basis at the leaf edge's own Gauss points
// (DGBoundaryCell::init, DGCell.cc:423-555); here each element with such faces
// integrates them itself (no atomics: each half's flux is formed by both of its
// sides, as the stage kernels form every face flux), after the stage kernel
// that left those faces out (code BC_HANG).  __host__ __device__, so the CPU
// emulation runs the same code.
#pragma once
#include <cmath>

#include "dg_kernels.h"
#include "dg_tables.h"

namespace dgp {

// the coefficient of v = dt R / |J| in the stage output and in the RK4
// accumulator of each stage update (dg_stage_tiled.h stage_value_hd)
DG_HD inline void stage_coeffs(int scheme, int stage, double& cout, double& cacc) {
  cacc = 0.0;
  switch (scheme * 4 + stage) {
    case 3 * 4 + 0: cout = 0.5; cacc = 1.0; break;
    case 3 * 4 + 1: cout = 0.5; cacc = 2.0; break;
    case 3 * 4 + 2: cout = 1.0; cacc = 2.0; break;
    case 3 * 4 + 3: cout = 1.0 / 6.0; break;
    case 2 * 4 + 0: cout = 1.0; break;
    case 2 * 4 + 1: cout = 0.25; break;
    case 2 * 4 + 2: cout = 2.0 / 3.0; break;
    case 1 * 4 + 0: cout = 1.0; break;
    case 1 * 4 + 1: cout = 0.5; break;
    default: cout = 1.0; break;  // forward Euler
  }
}

// element i of the list: its faces of half an edge, then the update
template <int P, class LAW>
DG_HD void hang_element(const HangArgs& a, int i) {
  constexpr int NP = np_tri(P), NQF = nqf_tri(P), NF = LAW::NF;
  const int E = a.E, e = a.elem[i];
  const double* tb = a.tab;
  double U[NF][NP], dR[NF][NP];
#pragma unroll
  for (int f = 0; f < NF; f++)
#pragma unroll
    for (int k = 0; k < NP; k++) {
      U[f][k] = a.Uin[(long)(f * NP + k) * E + e];
      dR[f][k] = 0.0;
    }
  const double m0 = a.el.mg[e], m1 = a.el.mg[(long)E + e], m2 = a.el.mg[2L * E + e], m3 = a.el.mg[3L * E + e];
  for (int r = a.rstart[i]; r < a.rstart[i + 1]; r++) {
    const int lo = a.rec[r * 4], nb = a.rec[r * 4 + 1], lnb = a.rec[r * 4 + 2], fl = a.rec[r * 4 + 3];
    const bool coarse = fl & 1, rev = fl & 2;
    const int h = (fl >> 2) & 1;
    // outward normal x length of the own face l: |J| J^-T nhat |fhat| with
    // nhat |fhat| of the reference edges (0,-1), (1,1), (-1,0)
    double nl0, nl1;
    if (lo == 0) {
      nl0 = -m2;
      nl1 = -m3;
    } else if (lo == 1) {
      nl0 = m0 + m2;
      nl1 = m1 + m3;
    } else {
      nl0 = -m0;
      nl1 = -m1;
    }
    const double len = dg_sqrt(nl0 * nl0 + nl1 * nl1);
    const double n[2] = {nl0 / len, nl1 / len};
    const double slen = coarse ? 0.5 * len : len;  // the half's length
    double V[NF][NP];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int k = 0; k < NP; k++) V[f][k] = a.Uin[(long)(f * NP + k) * E + nb];
    for (int q = 0; q < NQF; q++) {
      const int qq = rev ? NQF - 1 - q : q;
      // own basis row at the half's point q, the other side's at the same point
      const double* ro = coarse ? a.phih + ((lo * 2 + h) * NQF + q) * NP : tb + TriBlock::off_phif(P) + (lo * NQF + q) * NP;
      const double* rn = coarse ? tb + TriBlock::off_phif(P) + (lnb * NQF + qq) * NP : a.phih + ((lnb * 2 + h) * NQF + qq) * NP;
      double uL[NF], uR[NF], F[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int k = 0; k < NP; k++) {
          s0 += ro[k] * U[f][k];
          s1 += rn[k] * V[f][k];
        }
        uL[f] = s0;
        uR[f] = s1;
      }
      LAW::riemann(a.law, n, uL, uR, F);
      const double sw = slen * tb[TriBlock::off_wf(P) + q];
#pragma unroll
      for (int f = 0; f < NF; f++)
#pragma unroll
        for (int k = 0; k < NP; k++) dR[f][k] -= ro[k] * F[f] * sw;
    }
  }
  const bool residual_only = a.scheme == MODE_RESIDUAL;
  double cout = 0.0, cacc = 0.0;
  if (!residual_only) stage_coeffs(a.scheme, a.stage, cout, cacc);
  const double dtj = a.dt / a.el.detj[e];
#pragma unroll
  for (int f = 0; f < NF; f++)
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const long j = (long)(f * NP + k) * E + e;
      if (residual_only) {
        a.R[j] += dR[f][k];
      } else {
        const double dv = dR[f][k] * dtj;
        a.Uout[j] += cout * dv;
        if (cacc != 0.0) a.acc[j] += cacc * dv;
      }
    }
}

}  // namespace dgp
