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
DG_HD void stage_coeffs(int scheme, int stage, double& cout, double& cacc) {
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

// One workgroup of HANG_LANES work-items per element of the list (element i):
//   hangA  lanes = the element's rows: its coefficients -> LDS
//   hangB  lanes = (face half, point) pairs: both traces, the interior two-point
//          flux times the half's length and weight -> LDS
//   hangC  lanes = rows: R -= sum over the halves' points of phi_k F, then the
//          stage update's share
template <int P, class LAW>
struct HangBlock {
  static constexpr int NP = np_tri(P), NQF = nqf_tri(P), NF = LAW::NF, NROW = NF * NP;
  static_assert(NROW <= HANG_LANES && HANG_MAXREC * NQF <= HANG_LANES, "one lane per row and per point");
  struct Smem {
    double U[NROW];
    double F[HANG_MAXREC][NQF][NF];  // flux x length x weight
  };

  DG_HD static void hangA(const HangArgs& a, int i, int lane, Smem& s) {
    if (lane < NROW) s.U[lane] = a.Uin[(long)lane * a.E + a.elem[i]];
  }

  DG_HD static void hangB(const HangArgs& a, int i, int lane, Smem& s) {
    const int r0 = a.rstart[i], nrec = a.rstart[i + 1] - r0;
    const int rr = lane / NQF, q = lane - rr * NQF;
    if (rr >= nrec) return;
    const int E = a.E, e = a.elem[i], r = r0 + rr;
    const int lo = a.rec[r * 4], nb = a.rec[r * 4 + 1], lnb = a.rec[r * 4 + 2], fl = a.rec[r * 4 + 3];
    const bool coarse = fl & 1, rev = fl & 2;
    const int h = (fl >> 2) & 1;
    // outward normal x length of the own face lo: |J| J^-T nhat |fhat| with
    // nhat |fhat| of the reference edges (0,-1), (1,1), (-1,0)
    const double m0 = a.el.mg[e], m1 = a.el.mg[(long)E + e], m2 = a.el.mg[2L * E + e], m3 = a.el.mg[3L * E + e];
    const double nl0 = lo == 0 ? -m2 : lo == 1 ? m0 + m2 : -m0;
    const double nl1 = lo == 0 ? -m3 : lo == 1 ? m1 + m3 : -m1;
    const double len = dg_sqrt(nl0 * nl0 + nl1 * nl1);
    const double n[2] = {nl0 / len, nl1 / len};
    const double slen = coarse ? 0.5 * len : len;  // the half's length
    const int qq = rev ? NQF - 1 - q : q;
    // own basis row at the half's point q, the other side's at the same point
    const double* tb = a.tab;
    const double* ro = coarse ? a.phih + ((lo * 2 + h) * NQF + q) * NP : tb + TriBlock::off_phif(P) + (lo * NQF + q) * NP;
    const double* rn = coarse ? tb + TriBlock::off_phif(P) + (lnb * NQF + qq) * NP : a.phih + ((lnb * 2 + h) * NQF + qq) * NP;
    double uL[NF], uR[NF], F[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int k = 0; k < NP; k++) {
        s0 += ro[k] * s.U[f * NP + k];
        s1 += rn[k] * a.Uin[(long)(f * NP + k) * E + nb];
      }
      uL[f] = s0;
      uR[f] = s1;
    }
    LAW::riemann(a.law, n, uL, uR, F);
    const double sw = slen * tb[TriBlock::off_wf(P) + q];
#pragma unroll
    for (int f = 0; f < NF; f++) s.F[rr][q][f] = F[f] * sw;
  }

  DG_HD static void hangC(const HangArgs& a, int i, int lane, const Smem& s) {
    if (lane >= NROW) return;
    const int f = lane / NP, k = lane - f * NP;
    const int r0 = a.rstart[i], nrec = a.rstart[i + 1] - r0;
    const int E = a.E, e = a.elem[i];
    const double* tb = a.tab;
    double dR = 0.0;
    for (int rr = 0; rr < nrec; rr++) {
      const int r = r0 + rr;
      const int lo = a.rec[r * 4], fl = a.rec[r * 4 + 3];
      const bool coarse = fl & 1;
      const int h = (fl >> 2) & 1;
      for (int q = 0; q < NQF; q++) {
        const double phi = coarse ? a.phih[((lo * 2 + h) * NQF + q) * NP + k] : tb[TriBlock::off_phif(P) + (lo * NQF + q) * NP + k];
        dR -= phi * s.F[rr][q][f];
      }
    }
    const long j = (long)lane * E + e;
    if (a.scheme == MODE_RESIDUAL) {
      a.R[j] += dR;
    } else {
      double cout, cacc;
      stage_coeffs(a.scheme, a.stage, cout, cacc);
      const double dv = dR * (a.dt / a.el.detj[e]);
      a.Uout[j] += cout * dv;
      if (cacc != 0.0) a.acc[j] += cacc * dv;
    }
  }
};

}  // namespace dgp
