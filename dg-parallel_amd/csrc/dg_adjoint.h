// Derivatives of the triangle residual for the discrete adjoint of the
// unlimited residual and RK4 step (no counterpart in the reference):
//
//   TiledVjp   w = A(U)^T z, A = dR/dU, as tiled phases (64 elements per
//              workgroup) mirroring the residual's: the volume term transposes
//              "interpolate -> flux -> test" into "test^T -> flux^T -> interp^T",
//              and each face point adds J^T (zbar_own - zbar_nb) with
//              J = dFhat(n, u_own, u_nb)/du_own, using Fhat(-n, b, a) = -Fhat(n, a, b)
//              for the neighbour's evaluation of the shared face.  Flux Jacobians
//              come from forward-mode duals of the same flux templates.  An
//              optional epilogue fuses one reverse RK4 stage.
//   jvp_element  dR = A(U) dU for one element (tangent-linear model), forward
//              duals through the same residual arithmetic; used for the
//              tangent-linear / adjoint dot-product check.
// Both are __host__ __device__ so the CPU emulation runs the same code.
#pragma once

#include "dg_dual.h"
#include "dg_kernels.h"
#include "dg_stage_tiled.h"
#include "dg_tables.h"

namespace dgp {

enum { ADJ_PLAIN = 0, ADJ_RK4 = 1 };

struct AdjArgs {
  int E, P;
  const double* U;    // state at which the derivative is taken (SoA)
  const double* Z;    // vjp: seed z; jvp: tangent dU (SoA)
  double* W;          // vjp (plain): A^T z; jvp: A dU (SoA)
  const double* tab;
  TriElemTable el;
  LawParams law;
  int mode;           // ADJ_PLAIN or ADJ_RK4
  int stage;          // reverse RK4 stage 3, 2, 1, 0
  double dt;
  const double* lam;  // adjoint of U^{n+1}
  double* mu;         // adjoint of U^n (accumulated)
  double* znext;      // dt/|J| * kbar of the next (earlier) stage
};

// ND: tangent directions carried per flux evaluation (a divisor of NF).  The
// flux Jacobian at a point is built in NF/ND passes of Dual<double, ND>: fewer
// directions per pass re-evaluate the flux more often but keep fewer registers
// live (the Euler flux with 4 directions needs more than 256 VGPRs).
template <int P, class LAW, int ND = LAW::NF>
struct TiledVjp {
  static constexpr int NP = np_tri(P), NQ = nqv_tri(P), NQF = nqf_tri(P), NF = LAW::NF;
  static_assert(NF % ND == 0, "ND must divide NF");
  static constexpr int T = TILE_ELEMS;
  static constexpr int NROW = NF * NP;
  static constexpr int RPW = (NROW + 3) / 4;
  static constexpr int FW = NQF * NF;  // one face: points x fields
  static constexpr int WORK = NQ * NF > 2 * FW ? NQ * NF : 2 * FW;
  using D = Dual<double, ND>;

  // the face term runs one local face at a time (C1 -> C2 -> lift, per face), so
  // the work area holds one face's traces: 67 KB at P3 Euler, two workgroups per CU
  struct Smem {
    double Us[NROW][T];
    double Zs[NROW][T];
    double Ms[4][T];
    int Ls[3][T];
    double W[WORK][T];
  };
  struct Regs {
    double R[RPW];
  };

  DG_HD static void phaseA(const AdjArgs& a, Smem& s, int tid, int tile) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, e = tile * T + t;
    const bool valid = e < E;
    const int es = valid ? e : 0;  // padding lanes load element 0 (in bounds) and discard it
    // every load is issued, unconditionally, before the LDS writes use any of them
    double u[RPW], z[RPW];
#pragma unroll
    for (int j = 0; j < RPW; j++) {
      const int r = wave + 4 * j;
      u[j] = r < NROW ? a.U[(long)r * E + es] : 0.0;
      z[j] = r < NROW ? a.Z[(long)r * E + es] : 0.0;
    }
    const double mg = a.el.mg[(long)(wave & 3) * E + es];
    const int lk = a.el.link[(long)(wave < 3 ? wave : 0) * E + es];
#pragma unroll
    for (int j = 0; j < RPW; j++) {
      const int r = wave + 4 * j;
      if (r < NROW) {
        s.Us[r][t] = valid ? u[j] : (r == 0 ? 1.0 : 0.0);
        s.Zs[r][t] = valid ? z[j] : 0.0;
      }
    }
    if (wave < 4) s.Ms[wave][t] = valid ? mg : (wave == 0 || wave == 3 ? 1.0 : 0.0);
    if (wave < 3) s.Ls[wave][t] = valid ? lk : -BC_FARFIELD;
  }

  // volume: ubar_q = w_q sum_d (dF_d/du)^T Y_d, Y_d = sum_r mg[r][d] sum_k dphi_k^r z_k
  DG_HD static void phaseB(const AdjArgs& a, Smem& s, int tid) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = a.tab;
    for (int q = wave; q < NQ; q += 4) {
      double uq[NF], Y0[NF], Y1[NF];
      const double m00 = s.Ms[0][t], m01 = s.Ms[1][t], m10 = s.Ms[2][t], m11 = s.Ms[3][t];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        double acc = 0, z0 = 0, z1 = 0;
#pragma unroll
        for (int k = 0; k < NP; k++) {
          acc += tb[TriBlock::off_phi(P) + q * NP + k] * s.Us[f * NP + k][t];
          z0 += tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 0] * s.Zs[f * NP + k][t];
          z1 += tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 1] * s.Zs[f * NP + k][t];
        }
        uq[f] = acc;
        Y0[f] = m00 * z0 + m10 * z1;
        Y1[f] = m01 * z0 + m11 * z1;
      }
      const double w = tb[TriBlock::off_wv(P) + q];
#pragma unroll 1
      for (int j0 = 0; j0 < NF; j0 += ND) {
        D u[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          u[f] = D(uq[f]);
#pragma unroll
          for (int i = 0; i < ND; i++)
            if (f == j0 + i) u[f].d[i] = 1.0;
        }
        D F[NF][2];
        LAW::flux(a.law, u, F);
#pragma unroll
        for (int i = 0; i < ND; i++) {
          double ub = 0.0;
#pragma unroll
          for (int f = 0; f < NF; f++) ub += Y0[f] * F[f][0].d[i] + Y1[f] * F[f][1].d[i];
          s.W[q * NF + j0 + i][t] = w * ub;
        }
      }
    }
  }

  DG_HD static void phaseD1(const AdjArgs& a, Smem& s, int tid, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = a.tab;
#pragma unroll
    for (int j = 0; j < RPW; j++) {
      const int row = wave + 4 * j;
      double acc = 0.0;
      if (row < NROW) {
        const int f = row / NP, k = row - f * NP;
        for (int q = 0; q < NQ; q++) acc += tb[TriBlock::off_phi(P) + q * NP + k] * s.W[q * NF + f][t];
      }
      g.R[j] = acc;
    }
  }

  template <int LB, bool SAME>
  DG_HD static void nb_trace2(const AdjArgs& a, int nb, int f, double* ou, double* oz) {
    constexpr int OFF = TriBlock::off_phif(P);
    const double* tb = a.tab;
    double vu[NP], vz[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
      vu[k] = a.U[(long)(f * NP + k) * a.E + nb];
      vz[k] = a.Z[(long)(f * NP + k) * a.E + nb];
    }
#pragma unroll
    for (int q = 0; q < NQF; q++) {
      const int qq = SAME ? q : NQF - 1 - q;
      double au = 0.0, az = 0.0;
#pragma unroll
      for (int k = 0; k < NP; k++) {
        const double phi = tb[OFF + (LB * NQF + qq) * NP + k];
        au += phi * vu[k];
        az += phi * vz[k];
      }
      ou[q] = au;
      oz[q] = az;
    }
  }

  // face l: neighbour traces of the state (rows [0, FW)) and of the seed (rows [FW, 2FW))
  DG_HD static void phaseC1(const AdjArgs& a, Smem& s, int tid, int tile, int l) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const bool valid = tile * T + t < a.E;
    const int link = s.Ls[l][t];
    for (int f = wave; f < NF; f += 4) {
      double ou[NQF], oz[NQF];
      if (valid && link >= 0) {
        const int nb = link >> 3, lb = link & 3;
        const bool same = (link >> 2) & 1;
        if (!same) {
          if (lb == 0)
            nb_trace2<0, false>(a, nb, f, ou, oz);
          else if (lb == 1)
            nb_trace2<1, false>(a, nb, f, ou, oz);
          else
            nb_trace2<2, false>(a, nb, f, ou, oz);
        } else {
          if (lb == 0)
            nb_trace2<0, true>(a, nb, f, ou, oz);
          else if (lb == 1)
            nb_trace2<1, true>(a, nb, f, ou, oz);
          else
            nb_trace2<2, true>(a, nb, f, ou, oz);
        }
      } else {
#pragma unroll
        for (int q = 0; q < NQF; q++) ou[q] = oz[q] = 0.0;
      }
#pragma unroll
      for (int q = 0; q < NQF; q++) {
        s.W[q * NF + f][t] = ou[q];
        s.W[FW + q * NF + f][t] = oz[q];
      }
    }
  }

  // face l points: ubar = J^T (zbar_own - zbar_nb), written over the state traces
  // (each point's rows are read and rewritten by the one thread that owns it)
  DG_HD static void phaseC2(const AdjArgs& a, Smem& s, int tid, int l) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = a.tab;
    const int link = s.Ls[l][t];
    double nl0, nl1;
    if (l == 0) {
      nl0 = -s.Ms[2][t];
      nl1 = -s.Ms[3][t];
    } else if (l == 1) {
      nl0 = s.Ms[0][t] + s.Ms[2][t];
      nl1 = s.Ms[1][t] + s.Ms[3][t];
    } else {
      nl0 = -s.Ms[0][t];
      nl1 = -s.Ms[1][t];
    }
    const double len = dg_sqrt(nl0 * nl0 + nl1 * nl1);
    for (int q = wave; q < NQF; q += 4) {
      const D n[2] = {D(nl0 / len), D(nl1 / len)};
      const double sw = len * tb[TriBlock::off_wf(P) + q];
      double uLq[NF], uRq[NF], zb[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        double au = 0, az = 0;
#pragma unroll
        for (int k = 0; k < NP; k++) {
          const double phi = tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k];
          au += phi * s.Us[f * NP + k][t];
          az += phi * s.Zs[f * NP + k][t];
        }
        uLq[f] = au;
        uRq[f] = s.W[q * NF + f][t];
        const double znb = link >= 0 ? s.W[FW + q * NF + f][t] : 0.0;
        zb[f] = -sw * (az - znb);
      }
      // the point's rows are rewritten in place: its traces are in registers now
#pragma unroll 1
      for (int j0 = 0; j0 < NF; j0 += ND) {
        D uL[NF], uR[NF], Fh[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          uL[f] = D(uLq[f]);
          uR[f] = D(uRq[f]);
#pragma unroll
          for (int i = 0; i < ND; i++)
            if (f == j0 + i) uL[f].d[i] = 1.0;
        }
        face_flux<LAW>(a.law, link, n, uL, uR, Fh);
#pragma unroll
        for (int i = 0; i < ND; i++) {
          double acc = 0.0;
#pragma unroll
          for (int f = 0; f < NF; f++) acc += Fh[f].d[i] * zb[f];
          s.W[q * NF + j0 + i][t] = acc;
        }
      }
    }
  }

  // face l: lift transpose into the register rows
  DG_HD static void phaseLift(const AdjArgs& a, Smem& s, int tid, int l, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = a.tab;
#pragma unroll
    for (int j = 0; j < RPW; j++) {
      const int row = wave + 4 * j;
      if (row < NROW) {
        const int f = row / NP, k = row - f * NP;
        double acc = g.R[j];
#pragma unroll
        for (int q = 0; q < NQF; q++) acc += tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k] * s.W[q * NF + f][t];
        g.R[j] = acc;
      }
    }
  }

  // output: plain w = A^T z, or one reverse RK4 stage
  DG_HD static void phaseOut(const AdjArgs& a, int tid, int tile, const Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, e = tile * T + t;
    if (e >= E) return;
    const double dtj = a.dt / a.el.detj[e];
    // the rows' lam and mu are loaded before any is used
    const bool rk = a.mode != ADJ_PLAIN;
    double lamv[RPW], muv[RPW];
#pragma unroll
    for (int j = 0; j < RPW; j++) {
      const int row = wave + 4 * j;
      const long i = (long)(row < NROW ? row : 0) * E + e;
      lamv[j] = rk ? a.lam[i] : 0.0;
      muv[j] = rk && a.stage < 3 ? a.mu[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < RPW; j++) {
      const int row = wave + 4 * j;
      if (row < NROW) {
        const double acc = g.R[j];
        const long i = (long)row * E + e;
        if (!rk) {
          a.W[i] = acc;
        } else {
          // reverse RK4 (forward stages TimeIntegrator.cc:81-275): with ubar_s the
          // adjoint through stage s, mu = lam + ubar_3 + ubar_2 + ubar_1 + ubar_0 and
          // kbar_2 = lam/3 + ubar_3, kbar_1 = lam/3 + ubar_2/2, kbar_0 = lam/6 + ubar_1/2
          const double lam = lamv[j];
          if (a.stage == 3) {
            a.mu[i] = lam + acc;
            a.znext[i] = dtj * (lam * (1.0 / 3.0) + acc);
          } else if (a.stage == 2) {
            a.mu[i] = muv[j] + acc;
            a.znext[i] = dtj * (lam * (1.0 / 3.0) + 0.5 * acc);
          } else if (a.stage == 1) {
            a.mu[i] = muv[j] + acc;
            a.znext[i] = dtj * (lam * (1.0 / 6.0) + 0.5 * acc);
          } else {
            a.mu[i] = muv[j] + acc;
          }
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// tangent-linear residual of one element: dR = A(U) dU
template <int P, class LAW>
DG_HD void jvp_element(const AdjArgs& a, int e) {
  constexpr int NP = np_tri(P), NQ = nqv_tri(P), NQF = nqf_tri(P), NF = LAW::NF;
  using D = Dual<double, 1>;
  const int E = a.E;
  const double* tb = a.tab;
  D U[NF][NP], R[NF][NP];
  for (int f = 0; f < NF; f++)
    for (int k = 0; k < NP; k++) {
      const long i = (long)(f * NP + k) * E + e;
      U[f][k] = D(a.U[i]);
      U[f][k].d[0] = a.Z[i];
      R[f][k] = D(0.0);
    }
  double mg[2][2];
  for (int r = 0; r < 2; r++)
    for (int d = 0; d < 2; d++) mg[r][d] = a.el.mg[(long)(r * 2 + d) * E + e];
  for (int q = 0; q < NQ; q++) {
    D u[NF];
    for (int f = 0; f < NF; f++) {
      D s(0.0);
      for (int k = 0; k < NP; k++) s += tb[TriBlock::off_phi(P) + q * NP + k] * U[f][k];
      u[f] = s;
    }
    D F[NF][2];
    LAW::flux(a.law, u, F);
    const double w = tb[TriBlock::off_wv(P) + q];
    for (int f = 0; f < NF; f++) {
      D G0 = w * (mg[0][0] * F[f][0] + mg[0][1] * F[f][1]);
      D G1 = w * (mg[1][0] * F[f][0] + mg[1][1] * F[f][1]);
      for (int k = 0; k < NP; k++)
        R[f][k] += tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 0] * G0 +
                   tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 1] * G1;
    }
  }
  for (int l = 0; l < 3; l++) {
    const int link = a.el.link[(long)l * E + e];
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
    const double len = dg_sqrt(nl0 * nl0 + nl1 * nl1);
    const D n[2] = {D(nl0 / len), D(nl1 / len)};
    for (int q = 0; q < NQF; q++) {
      D uL[NF], uR[NF], Fh[NF];
      for (int f = 0; f < NF; f++) {
        D s(0.0);
        for (int k = 0; k < NP; k++) s += tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k] * U[f][k];
        uL[f] = s;
      }
      if (link >= 0) {
        const int nb = link >> 3, lb = link & 3;
        const bool same = (link >> 2) & 1;
        const int qq = same ? q : NQF - 1 - q;
        for (int f = 0; f < NF; f++) {
          D s(0.0);
          for (int k = 0; k < NP; k++) {
            const long i = (long)(f * NP + k) * E + nb;
            D v(a.U[i]);
            v.d[0] = a.Z[i];
            s += tb[TriBlock::off_phif(P) + (lb * NQF + qq) * NP + k] * v;
          }
          uR[f] = s;
        }
      }
      face_flux<LAW>(a.law, link, n, uL, uR, Fh);
      const double sw = len * tb[TriBlock::off_wf(P) + q];
      for (int f = 0; f < NF; f++)
        for (int k = 0; k < NP; k++) R[f][k] -= (sw * tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k]) * Fh[f];
    }
  }
  for (int f = 0; f < NF; f++)
    for (int k = 0; k < NP; k++) a.W[(long)(f * NP + k) * E + e] = R[f][k].d[0];
}

}  // namespace dgp
