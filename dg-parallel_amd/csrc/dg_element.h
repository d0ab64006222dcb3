// The element-per-work-item triangle stage (tri_stage_kernel in
// dg_kernels.hip): each work-item owns one element, keeps its modal
// coefficients and residual in registers, and
//   1. volume term (DGCell.cc:184-249): at every volume point interpolates the
//      state, evaluates the physical flux, maps it with |J| J^-1 (one 2x2 per
//      element) and contracts with the reference basis gradients;
//   2. face term (DGCell.cc:639-732): for each of its three faces evaluates its
//      own trace and the neighbour's trace (neighbour coefficients read straight
//      from the stage-input array, which no kernel writes during the stage),
//      the numerical flux, and lifts it with its own trace basis.  Each face
//      flux is evaluated by both neighbours: no face buffer, no atomics, and
//      the result does not depend on scheduling;
//   3. stage update (TimeIntegrator.cc:81-275, 279-438, 838-990): applies the
//      inverse mass 1/|J| and the RK stage formula, reading/writing the FP64
//      step-start state and accumulator and writing the next stage input.
// Tables (basis values at quadrature points) are addressed with compile-time
// offsets from a kernel-argument pointer, so they are wave-uniform scalar loads.
// Elements are ordered on the host so that a wavefront shares the neighbour's
// local-face index on each face; the switch below then runs one path per wave.
// __host__ __device__, so the CPU emulation runs the same code.
#pragma once

#include "dg_kernels.h"
#include "dg_tables.h"

namespace dgp {

template <int P, class LAW, typename CT, typename IT, int LB, bool SAME>
DG_HD void nb_traces(const IT* __restrict__ Uin, int E, int nb, const CT* __restrict__ tb,
                                          CT (&uR)[nqf_tri(P)][LAW::NF]) {
  constexpr int NP = np_tri(P), NQF = nqf_tri(P), NF = LAW::NF;
  constexpr int OFF = TriBlock::off_phif(P);
#pragma unroll
  for (int f = 0; f < NF; f++) {
    CT v[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) v[k] = CT(Uin[(long)(f * NP + k) * E + nb]);
#pragma unroll
    for (int q = 0; q < NQF; q++) {
      const int qq = SAME ? q : NQF - 1 - q;
      CT s = CT(0);
#pragma unroll
      for (int k = 0; k < NP; k++) s += tb[OFF + (LB * NQF + qq) * NP + k] * v[k];
      uR[q][f] = s;
    }
  }
}

template <int P, class LAW, typename CT, typename IT>
DG_HD void nb_traces_any(const IT* __restrict__ Uin, int E, int link, const CT* __restrict__ tb,
                                              CT (&uR)[nqf_tri(P)][LAW::NF]) {
  int nb = link >> 3;
  int lb = link & 3;
  bool same = (link >> 2) & 1;
  if (!same) {
    if (lb == 0)
      nb_traces<P, LAW, CT, IT, 0, false>(Uin, E, nb, tb, uR);
    else if (lb == 1)
      nb_traces<P, LAW, CT, IT, 1, false>(Uin, E, nb, tb, uR);
    else
      nb_traces<P, LAW, CT, IT, 2, false>(Uin, E, nb, tb, uR);
  } else {
    if (lb == 0)
      nb_traces<P, LAW, CT, IT, 0, true>(Uin, E, nb, tb, uR);
    else if (lb == 1)
      nb_traces<P, LAW, CT, IT, 1, true>(Uin, E, nb, tb, uR);
    else
      nb_traces<P, LAW, CT, IT, 2, true>(Uin, E, nb, tb, uR);
  }
}

// The whole stage of element e (one work-item): volume term, face terms, stage
// update; rsum gets the element's share of the step residual norm.
template <int P, class LAW, typename CT, typename IT, typename OT>
DG_HD void element_stage(const StageArgs& a, int e, double& rsum) {
  constexpr int NP = np_tri(P), NQ = nqv_tri(P), NQF = nqf_tri(P), NF = LAW::NF;
  const int E = a.E;
  {
    const CT* __restrict__ tb = sizeof(CT) == 8 ? (const CT*)a.tab : (const CT*)a.tabf;
    const IT* __restrict__ Uin = (const IT*)a.Uin;
    const LawParams& law = a.law;

    CT U[NF][NP];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int k = 0; k < NP; k++) U[f][k] = CT(Uin[(long)(f * NP + k) * E + e]);
    CT mg[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
      for (int d = 0; d < 2; d++) mg[r][d] = CT(a.el.mg[(long)(r * 2 + d) * E + e]);

    CT R[NF][NP];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int k = 0; k < NP; k++) R[f][k] = CT(0);

    // ---- volume term -------------------------------------------------------
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      CT u[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        CT s = CT(0);
#pragma unroll
        for (int k = 0; k < NP; k++) s += tb[TriBlock::off_phi(P) + q * NP + k] * U[f][k];
        u[f] = s;
      }
      CT F[NF][2];
      LAW::flux(law, u, F);
      const CT w = tb[TriBlock::off_wv(P) + q];
      CT G[NF][2];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        G[f][0] = w * (mg[0][0] * F[f][0] + mg[0][1] * F[f][1]);
        G[f][1] = w * (mg[1][0] * F[f][0] + mg[1][1] * F[f][1]);
      }
#pragma unroll
      for (int k = 0; k < NP; k++) {
        const CT g0 = tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 0];
        const CT g1 = tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 1];
#pragma unroll
        for (int f = 0; f < NF; f++) R[f][k] += g0 * G[f][0] + g1 * G[f][1];
      }
    }

    // ---- face terms ----------------------------------------------------------
#pragma unroll
    for (int l = 0; l < 3; l++) {
      const int link = a.el.link[(long)l * E + e];
      // outward normal times edge length: |J| J^-T nhat |fhat| (nhat |fhat| of the
      // reference edges: (0,-1), (1,1), (-1,0))
      CT nl0, nl1;
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
      const CT len = dg_sqrt(nl0 * nl0 + nl1 * nl1);
      const CT n[2] = {nl0 / len, nl1 / len};
      CT uR[NQF][NF];
      if (link >= 0) nb_traces_any<P, LAW, CT, IT>(Uin, E, link, tb, uR);
#pragma unroll
      for (int q = 0; q < NQF; q++) {
        CT uL[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          CT s = CT(0);
#pragma unroll
          for (int k = 0; k < NP; k++) s += tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k] * U[f][k];
          uL[f] = s;
        }
        CT Fh[NF];
        face_flux<LAW>(law, link, n, uL, uR[q], Fh);
        const CT sw = len * tb[TriBlock::off_wf(P) + q];  // |J_face| w = (len/2) w
#pragma unroll
        for (int f = 0; f < NF; f++) {
          const CT fs = Fh[f] * sw;
#pragma unroll
          for (int k = 0; k < NP; k++) R[f][k] -= fs * tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k];
        }
      }
    }

    // ---- stage update ------------------------------------------------------
    if (a.scheme == MODE_RESIDUAL) {
#pragma unroll
      for (int f = 0; f < NF; f++)
#pragma unroll
        for (int k = 0; k < NP; k++) a.R[(long)(f * NP + k) * E + e] = double(R[f][k]);
    } else {
      const double dtj = a.dt / a.el.detj[e];
      OT* __restrict__ Uout = (OT*)a.Uout;
      const double* __restrict__ Un = a.Un;
      double* __restrict__ acc = a.acc;
      const int sc = a.scheme * 4 + a.stage;
#pragma unroll
      for (int f = 0; f < NF; f++)
#pragma unroll
        for (int k = 0; k < NP; k++) {
          const long i = (long)(f * NP + k) * E + e;
          const double v = double(R[f][k]) * dtj;
          double o;
          switch (sc) {
            case 3 * 4 + 0: {  // RK4 stage 1 (TimeIntegrator.cc:101-137)
              acc[i] = v;
              o = Un[i] + 0.5 * v;
              break;
            }
            case 3 * 4 + 1: {  // RK4 stage 2 (:149-181)
              acc[i] += 2. * v;
              o = Un[i] + 0.5 * v;
              break;
            }
            case 3 * 4 + 2: {  // RK4 stage 3 (:192-222)
              acc[i] += 2. * v;
              o = Un[i] + v;
              break;
            }
            case 3 * 4 + 3: {  // RK4 stage 4 (:235-270)
              const double s = acc[i] + v;
              o = Un[i] + s / 6.0;
              rsum += s * s / 36.0;
              break;
            }
            case 2 * 4 + 0:  // SSP-RK3 stage 1 (:850-880)
              o = Un[i] + v;
              break;
            case 2 * 4 + 1:  // SSP-RK3 stage 2 (:889-925)
              o = 0.25 * (double(Uin[i]) + v + 3. * Un[i]);
              break;
            case 2 * 4 + 2: {  // SSP-RK3 stage 3 (:934-987)
              o = (2.0 * (double(Uin[i]) + v) + Un[i]) / 3.0;
              const double d = o - Un[i];
              rsum += d * d;
              break;
            }
            case 1 * 4 + 0:  // SSP-RK2 stage 1 (parallel.cc:78-131)
              o = Un[i] + v;
              break;
            case 1 * 4 + 1: {  // SSP-RK2 stage 2 (parallel.cc:133-186)
              o = (double(Uin[i]) + v + Un[i]) * 0.5;
              const double d = o - Un[i];
              rsum += d * d;
              break;
            }
            default: {  // forward Euler (TimeIntegrator.cc:993-1048)
              o = Un[i] + v;
              const double d = Un[i] - o;
              rsum += d * d;
              break;
            }
          }
          Uout[i] = OT(o);
          if constexpr (sizeof(CT) == 4 && sizeof(OT) == 8) {
            if (a.Uout32) a.Uout32[i] = float(o);
          }
        }
    }
  }
}

}  // namespace dgp
