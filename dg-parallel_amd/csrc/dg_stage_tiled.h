// Tiled residual + RK stage for triangles, written as a sequence of phases
// separated by workgroup barriers.  Each phase is a __host__ __device__ function
// of (thread id, LDS tile), so the same code runs in the HIP kernel
// (dg_stage_tiled.hip) and, thread by thread, in a CPU emulation used by the
// tests.
//
// One workgroup (4 wavefronts) owns a tile of 64 elements, one element per
// lane.  Every phase maps a wave-uniform work index (quadrature point, face,
// output row) to the wave and the element to the lane, so basis tables are
// scalar loads and LDS accesses are lane-contiguous.
//   A  stage-input tile, |J| J^-1 and face links -> LDS
//   B  volume points (wave = point q): interpolate, physical flux, contravariant
//      flux G = w_q |J| J^-1 F -> LDS                      (DGCell.cc:193-236)
//   D1 output rows (wave = row f*NP+k): R = sum_q dphi_k(q) . G(q)   (registers)
//   C1 neighbour traces (wave = (face, field)): coefficients of the neighbour
//      read from the stage input in global memory (L2-resident), evaluated at
//      the face points in the neighbour's frame -> LDS
//   C2 face points (wave = (face, point)): own trace, numerical flux, weight
//      -> LDS                                               (DGCell.cc:647-705)
//   D2 output rows: R -= sum_{l,q} phi_k(l,q) Fhat(l,q)        (DGCell.cc:709-728)
//      then the stage update with the FP64 step-start state and accumulator
// Each face flux is computed by both of its elements (no face buffer, no
// atomics; results independent of scheduling).
#pragma once
#include <type_traits>

#include "dg_kernels.h"
#include "dg_laws.h"
#include "dg_tables.h"

namespace dgp {

constexpr int TILE_THREADS = 256;
constexpr int TILE_ELEMS = 64;

DG_HD int wave_uniform(int x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_readfirstlane(x);
#else
  return x;
#endif
}

// true when the predicate holds on every active lane of the wave (on the CPU
// emulation, which runs one lane at a time, the lane's own predicate)
DG_HD bool wave_all(bool pred) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __all(pred);
#else
  return pred;
#endif
}

// Divisions by 3, 6 and 36 are multiplications by the rounded reciprocals (one
// rounding more than the reference's division, far inside the 1e-12 parity
// bound; a full FP64 division is a 10-instruction sequence per value).
DG_HD double stage_value_hd(int sc, double v, double un, double us, double* acc, double& rsum) {
  constexpr double third = 1.0 / 3.0, sixth = 1.0 / 6.0, r36 = 1.0 / 36.0;
  double o;
  switch (sc) {
    case 3 * 4 + 0:  // RK4 (TimeIntegrator.cc:101-137)
      *acc = v;
      o = un + 0.5 * v;
      break;
    case 3 * 4 + 1:  // (:149-181)
      *acc += 2. * v;
      o = un + 0.5 * v;
      break;
    case 3 * 4 + 2:  // (:192-222)
      *acc += 2. * v;
      o = un + v;
      break;
    case 3 * 4 + 3: {  // (:235-270)
      const double s = *acc + v;
      o = un + s * sixth;
      rsum += s * s * r36;
      break;
    }
    case 2 * 4 + 0:  // SSP-RK3 (:850-880)
      o = un + v;
      break;
    case 2 * 4 + 1:  // (:889-925)
      o = 0.25 * (us + v + 3. * un);
      break;
    case 2 * 4 + 2: {  // (:934-987)
      o = (2.0 * (us + v) + un) * third;
      const double d = o - un;
      rsum += d * d;
      break;
    }
    case 1 * 4 + 0:  // SSP-RK2 (parallel.cc:78-131)
      o = un + v;
      break;
    case 1 * 4 + 1: {  // (parallel.cc:133-186)
      o = (us + v + un) * 0.5;
      const double d = o - un;
      rsum += d * d;
      break;
    }
    default: {  // forward Euler (TimeIntegrator.cc:993-1048)
      o = un + v;
      const double d = un - o;
      rsum += d * d;
      break;
    }
  }
  return o;
}

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// D += A B for one wave: v_mfma_f32_16x16x4_f32.  Lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and holds D[4 (l / 16) + r][l % 16], r = 0..3.
DG_HD mfma_f4 mfma_16x16x4(float a, float b, mfma_f4 c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
#else
  // the MFMA tile shape runs on the device only (no host instantiation calls it)
  __builtin_trap();
  return c;
#endif
}

struct MfmaAccOn {
  mfma_f4 a[4];  // column blocks of 16 elements of the wave's field
};
struct MfmaAccOff {};

// NW waves share a tile of 64 elements (NW = 4: 256 threads; NW = 8: 512
// threads, the same LDS and half the rows per thread); the tile-shape choice of
// the tuner.  MF (FP32 form, four fields, four waves): the test contraction
// (D1) and the face lift (D2) run on the matrix cores, each wave one field,
// rows = modes and columns = the tile's elements in four blocks of 16.
template <int P, class LAW, typename CT, typename IT, typename OT, int NW = 4, bool MF = false>
struct TiledStage {
  static constexpr int NP = np_tri(P), NQ = nqv_tri(P), NQF = nqf_tri(P), NF = LAW::NF;
  static constexpr int T = TILE_ELEMS;
  static constexpr int THREADS = NW * T;
  static constexpr int NROW = NF * NP;
  // LDS work rows: the volume points run in two passes of QH points, and each
  // face flux is written over the neighbour trace of its point
  static constexpr int QH = (NQ + 1) / 2;
  static constexpr int GW = QH * NF * 2;   // volume-flux rows of one pass
  static constexpr int FW = 3 * NQF * NF;  // trace, then face-flux rows
  static constexpr int WORK = GW > FW ? GW : FW;
  // output rows of a wave: with four or more fields the waves form FG field
  // groups (f = w % FG, + FG, ..) times KG mode groups (k = w / FG, + KG, ..);
  // a single field is split by modes only -- either way the rows of one field
  // share every LDS read
  static constexpr bool BY_FIELD = NF >= 4;
  static constexpr int FG = BY_FIELD ? (NW < 4 ? NW : 4) : 1;
  static constexpr int KG = NW / FG;
  static constexpr int FPW = (NF + FG - 1) / FG;
  static constexpr int KPW = (NP + KG - 1) / KG;
  static constexpr int QPW = (QH + NW - 1) / NW;
  static constexpr int CPW = (3 * NF + NW - 1) / NW;
  static constexpr int FPPW = (3 * NQF + NW - 1) / NW;

  struct Smem {
    CT Us[NROW][T];
    CT Ms[4][T];
    CT Ns[3][3][T];  // per face: unit outward normal (2) and edge length
    int Ls[3][T];
    union {
      CT W[WORK][T];
      double red[THREADS];  // the tile's residual-norm reduction, after D2
    };
  };

  // per-thread output rows of its own element, and the neighbour coefficients of
  // its face tasks, loaded in phase A so their latency hides behind B and D1
  // (a workgroup barrier waits for LDS and scalar loads only)
  static_assert(!MF || (NF == 4 && NW == 4 && sizeof(CT) == 4 && NP <= 16), "MFMA tile: FP32, 4 fields, 4 waves");
  struct Regs {
    CT R[FPW * KPW];
    double rsum;
    int es;  // the lane's element (0 on padding lanes)
    CT nv[CPW][NP];
    int lk[CPW];
    typename std::conditional<MF, MfmaAccOn, MfmaAccOff>::type m;
  };

  // the tile's first element and the end of the launch's element range
  DG_HD static int range_end(const StageArgs& a) { return a.e1 > 0 ? a.e1 : a.E; }

  DG_HD static const CT* table(const StageArgs& a) {
    return sizeof(CT) == 8 ? (const CT*)(const void*)a.tab : (const CT*)(const void*)a.tabf;
  }
  // a group of size one contributes no wave term, so with 4 waves the mode
  // index of an Euler row (and the field index of a scalar row) is a
  // compile-time constant and the basis tables are read at fixed offsets
  DG_HD static int row_f(int wave, int fi) { return (FG == 1 ? 0 : KG == 1 ? wave : wave % FG) + FG * fi; }
  DG_HD static int row_k(int wave, int ki) { return (KG == 1 ? 0 : FG == 1 ? wave : wave / FG) + KG * ki; }

  // ---- A: stage-input tile -> LDS (rows split over the waves); own element's
  //      geometry and links -> registers ----------------------------------------
  DG_HD static void phaseA(const StageArgs& a, Smem& s, int tid, int tile, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, e = a.e0 + tile * T + t;
    const bool valid = e < range_end(a);
    const int es = valid ? e : a.e0;  // padding lanes load the range's first element (in bounds) and discard it
    const IT* Uin = (const IT*)a.Uin;
    g.es = es;
    // Every global load of the phase is issued, unconditionally and in this
    // order, before any of them is used: the links (the neighbour addresses need
    // only them; the memory counter is in order), the tile's rows and |J| J^-1,
    // then C1's neighbour rows.  The LDS writes come last.
#pragma unroll
    for (int i = 0; i < FPW * KPW; i++) g.R[i] = CT(0);
    if constexpr (MF) {
#pragma unroll
      for (int cb = 0; cb < 4; cb++) g.m.a[cb] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    }
    int lk3[3];
#pragma unroll
    for (int l = 0; l < 3; l++) lk3[l] = a.el.link[(long)l * E + es];
    constexpr int RA = (NROW + NW - 1) / NW;
    CT u[RA];
#pragma unroll
    for (int j = 0; j < RA; j++) {
      const int r = wave + NW * j;
      u[j] = r < NROW ? CT(Uin[(long)r * E + es]) : CT(0);
    }
    CT m[4];
#pragma unroll
    for (int r = 0; r < 4; r++) m[r] = CT(a.el.mg[(long)r * E + es]);
#pragma unroll
    for (int j = 0; j < CPW; j++) {
      const int it = wave + NW * j;
      const int l = it < 3 * NF ? it / NF : 0, f = it < 3 * NF ? it - l * NF : 0;
      const int lk = valid ? lk3[l] : -BC_FARFIELD;
      g.lk[j] = lk;
      const int src = lk >= 0 ? (lk >> 3) : es;  // boundary faces read their own element and discard it
#pragma unroll
      for (int k = 0; k < NP; k++) g.nv[j][k] = CT(Uin[(long)(f * NP + k) * E + src]);
    }
#pragma unroll
    for (int j = 0; j < RA; j++) {
      const int r = wave + NW * j;
      if (r < NROW) s.Us[r][t] = valid ? u[j] : CT(r == 0 ? 1 : 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (!valid) m[r] = CT(r == 0 || r == 3 ? 1 : 0);
      if (r == wave) s.Ms[r][t] = m[r];
    }
    if (wave < 3) {
      const int l = wave;
      s.Ls[l][t] = valid ? lk3[l] : -BC_FARFIELD;
      // outward normal x edge length = |J| J^-T nhat |fhat|, with nhat |fhat| of
      // the reference edges (0,-1), (1,1), (-1,0)
      CT nl0, nl1;
      if (l == 0) {
        nl0 = -m[2];
        nl1 = -m[3];
      } else if (l == 1) {
        nl0 = m[0] + m[2];
        nl1 = m[1] + m[3];
      } else {
        nl0 = -m[0];
        nl1 = -m[1];
      }
      const CT len = dg_sqrt(nl0 * nl0 + nl1 * nl1);
      s.Ns[l][0][t] = nl0 / len;
      s.Ns[l][1][t] = nl1 / len;
      s.Ns[l][2][t] = len;
    }
  }

  // ---- B: volume flux at the quadrature points [q0, q0 + QH) ----------------------
  DG_HD static void phaseB(const StageArgs& a, Smem& s, int tid, const Regs& g, int q0) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const CT* tb = table(a);
#pragma unroll 1
    for (int j = 0; j < QPW; j++) {
      const int qi = wave + NW * j, q = q0 + qi;
      if (qi < QH && q < NQ) {
        CT u[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          CT acc = CT(0);
#pragma unroll
          for (int k = 0; k < NP; k++) acc += tb[TriBlock::off_phi(P) + q * NP + k] * s.Us[f * NP + k][t];
          u[f] = acc;
        }
        CT F[NF][2];
        if constexpr (point_vel<LAW>::value) {
          const CT v[2] = {CT(a.pvel[(long)(q * 2 + 0) * a.E + g.es]), CT(a.pvel[(long)(q * 2 + 1) * a.E + g.es])};
          LAW::flux_at(v, u, F);
        } else {
          LAW::flux(a.law, u, F);
        }
        const CT w = tb[TriBlock::off_wv(P) + q];
        const CT m00 = s.Ms[0][t], m01 = s.Ms[1][t], m10 = s.Ms[2][t], m11 = s.Ms[3][t];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          s.W[(qi * NF + f) * 2 + 0][t] = w * (m00 * F[f][0] + m01 * F[f][1]);
          s.W[(qi * NF + f) * 2 + 1][t] = w * (m10 * F[f][0] + m11 * F[f][1]);
        }
      }
    }
  }

  // ---- D1: contraction of points [q0, q0 + QH) with the basis gradients ------
  DG_HD static void phaseD1(const StageArgs& a, Smem& s, int tid, Regs& g, int q0) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const CT* tb = table(a);
    if constexpr (MF) {
      // R[i][e] += sum over (point, direction) kk of dphi[q][i][d] G[q][f][d][e],
      // four kk per MFMA step; lane: column (element) l % 16, step row l / 16
      constexpr int KK = 2 * QH, KS = (KK + 3) / 4;
      const int col = t & 15, kr = t >> 4, f = wave;
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int kk = 4 * ks + kr, qi = kk >> 1, d = kk & 1, q = q0 + qi;
        const bool ok = kk < KK && q < NQ;
        const float A = ok && col < NP ? tb[TriBlock::off_dphi(P) + (q * NP + col) * 2 + d] : 0.f;
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
          const float B = ok ? s.W[(qi * NF + f) * 2 + d][16 * cb + col] : 0.f;
          g.m.a[cb] = mfma_16x16x4(A, B, g.m.a[cb]);
        }
      }
      return;
    }
    const int q1 = q0 + QH < NQ ? q0 + QH : NQ;
#pragma unroll
    for (int fi = 0; fi < FPW; fi++) {
      const int f = row_f(wave, fi);
      if (f < NF) {
        for (int q = q0; q < q1; q++) {
          const CT G0 = s.W[((q - q0) * NF + f) * 2 + 0][t];
          const CT G1 = s.W[((q - q0) * NF + f) * 2 + 1][t];
#pragma unroll
          for (int ki = 0; ki < KPW; ki++) {
            const int k = row_k(wave, ki);
            if (k < NP) {
              g.R[fi * KPW + ki] += tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 0] * G0;
              g.R[fi * KPW + ki] += tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 1] * G1;
            }
          }
        }
      }
    }
  }

  // ---- C1: neighbour traces ---------------------------------------------------------
  // the neighbour's coefficients of one field evaluated at this face's points,
  // in the neighbour's local-face frame
  template <int LB, bool SAME>
  DG_HD static void nb_combine(const CT* tb, const CT* v, CT* out) {
    constexpr int OFF = TriBlock::off_phif(P);
#pragma unroll
    for (int q = 0; q < NQF; q++) {
      const int qq = SAME ? q : NQF - 1 - q;
      CT acc = CT(0);
#pragma unroll
      for (int k = 0; k < NP; k++) acc += tb[OFF + (LB * NQF + qq) * NP + k] * v[k];
      out[q] = acc;
    }
  }

  DG_HD static void phaseC1(const StageArgs& a, Smem& s, int tid, int tile, const Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int e = a.e0 + tile * T + t;
    const bool valid = e < range_end(a);
    const CT* tb = table(a);
    const auto& v = g.nv;  // loaded in phase A
#pragma unroll
    for (int j = 0; j < CPW; j++) {
      const int it = wave + NW * j;
      if (it < 3 * NF) {
        const int l = it / NF, f = it - l * NF;
        const int link = g.lk[j];
        CT out[NQF];
        if (valid && link >= 0) {
          const int lb = link & 3;
          const bool same = (link >> 2) & 1;
          if (!same) {
            if (lb == 0)
              nb_combine<0, false>(tb, v[j], out);
            else if (lb == 1)
              nb_combine<1, false>(tb, v[j], out);
            else
              nb_combine<2, false>(tb, v[j], out);
          } else {
            if (lb == 0)
              nb_combine<0, true>(tb, v[j], out);
            else if (lb == 1)
              nb_combine<1, true>(tb, v[j], out);
            else
              nb_combine<2, true>(tb, v[j], out);
          }
        } else {
#pragma unroll
          for (int q = 0; q < NQF; q++) out[q] = CT(0);
        }
#pragma unroll
        for (int q = 0; q < NQF; q++) s.W[(l * NQF + q) * NF + f][t] = out[q];
      }
    }
  }

  // ---- C2: numerical flux at the face points ----------------------------------
  DG_HD static void phaseC2(const StageArgs& a, Smem& s, int tid, const Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const CT* tb = table(a);
#pragma unroll 1
    for (int j = 0; j < FPPW; j++) {
      const int it = wave + NW * j;
      if (it < 3 * NQF) {
        const int l = it / NQF, q = it - l * NQF;
        // a face of a mesh with hanging nodes adds nothing here (the hanging-face
        // pass integrates its halves): the flux of the own trace on both sides,
        // weight zero
        const bool hang = s.Ls[l][t] == -BC_HANG;
        const int link = hang ? 0 : s.Ls[l][t];
        const CT n[2] = {s.Ns[l][0][t], s.Ns[l][1][t]};
        const CT len = hang ? CT(0) : s.Ns[l][2][t];
        CT uL[NF], uR[NF], Fh[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          CT acc = CT(0);
#pragma unroll
          for (int k = 0; k < NP; k++) acc += tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k] * s.Us[f * NP + k][t];
          uL[f] = acc;
          uR[f] = hang ? acc : s.W[(l * NQF + q) * NF + f][t];
        }
        // boundary faces: the law's ghost state, then the same two-point flux (one
        // inlined flux for interior and boundary lanes alike)
        if constexpr (point_vel<LAW>::value) {
          const long r = (long)(NQ + l * NQF + q) * 2;
          const CT v[2] = {CT(a.pvel[r * a.E + g.es]), CT(a.pvel[(r + 1) * a.E + g.es])};
          if (link < 0) LAW::ghost(a.law, -link, n, uL, uR);
          LAW::riemann_at(v, n, uL, uR, Fh);
        } else {
          face_flux<LAW>(a.law, link, n, uL, uR, Fh);
        }
        // |J_face| w_q = (len / 2) w_q; the table holds w_q / 2
        const CT sw = len * tb[TriBlock::off_wf(P) + q];
#pragma unroll
        for (int f = 0; f < NF; f++) s.W[(l * NQF + q) * NF + f][t] = Fh[f] * sw;  // over the neighbour trace
      }
    }
  }

  // ---- MF: the face lift on the matrix cores, then the rows back to their
  //      elements through LDS (after a barrier: W holds the face fluxes) ------
  DG_HD static void phaseLiftMF(const StageArgs& a, const Smem& s, int tid, Regs& g) {
    if constexpr (MF) {
      const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
      const CT* tb = table(a);
      constexpr int KK = 3 * NQF, KS = (KK + 3) / 4;
      const int col = t & 15, kr = t >> 4, f = wave;
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int kk = 4 * ks + kr;
        const bool ok = kk < KK;
        const float A = ok && col < NP ? -tb[TriBlock::off_phif(P) + kk * NP + col] : 0.f;
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
          const float B = ok ? s.W[kk * NF + f][16 * cb + col] : 0.f;
          g.m.a[cb] = mfma_16x16x4(A, B, g.m.a[cb]);
        }
      }
    }
  }
  DG_HD static void phaseStoreMF(Smem& s, int tid, const Regs& g) {
    if constexpr (MF) {
      const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
      const int col = t & 15, kr = t >> 4, f = wave;
#pragma unroll
      for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = 4 * kr + r;
          if (i < NP) s.W[f * NP + i][16 * cb + col] = g.m.a[cb][r];
        }
    }
  }

  // ---- D2 + stage update ----------------------------------------------------
  DG_HD static void phaseD2(const StageArgs& a, Smem& s, int tid, int tile, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, e = a.e0 + tile * T + t;
    const bool valid = e < range_end(a);
    const int es = valid ? e : a.e0;
    const CT* tb = table(a);
    const IT* Uin = (const IT*)a.Uin;
    const int sc = a.scheme * 4 + a.stage;
    const bool residual_only = a.scheme == MODE_RESIDUAL;
    const bool lts = a.scheme == SCHEME_LTS;
    const bool reads_input = !lts && (sc == 2 * 4 + 1 || sc == 2 * 4 + 2 || sc == 1 * 4 + 1);
    const bool reads_acc = lts ? !(a.stage & LTS_RESET) : (sc == 3 * 4 + 1 || sc == 3 * 4 + 2 || sc == 3 * 4 + 3);
    const double dtj = valid ? a.dt / a.el.detj[e] : 0.0;
    g.rsum = 0.0;
#pragma unroll
    for (int fi = 0; fi < FPW; fi++) {
      const int f = row_f(wave, fi);
      if (f < NF) {
        // issue the step-start / accumulator loads first; the lift below hides them
        double un[KPW], ac[KPW], us[KPW];
#pragma unroll
        for (int ki = 0; ki < KPW; ki++) {
          const int k = row_k(wave, ki);
          const long i = (long)(f * NP + (k < NP ? k : 0)) * E + es;
          un[ki] = residual_only ? 0.0 : a.Un[i];
          ac[ki] = reads_acc ? a.acc[i] : 0.0;
          us[ki] = reads_input ? double(Uin[i]) : 0.0;
        }
        if constexpr (MF) {
          // the rows of volume and face terms, stored by phaseStoreMF
#pragma unroll
          for (int ki = 0; ki < KPW; ki++) {
            const int k = row_k(wave, ki);
            if (k < NP) g.R[fi * KPW + ki] = s.W[f * NP + k][t];
          }
        } else {
          // one face's points per unrolled block: unrolling all 3*NQF points hoists
          // every table row into SGPRs and spills them to VGPR lanes
#pragma unroll NQF
          for (int lq = 0; lq < 3 * NQF; lq++) {
            const CT Fq = s.W[lq * NF + f][t];
#pragma unroll
            for (int ki = 0; ki < KPW; ki++) {
              const int k = row_k(wave, ki);
              if (k < NP) g.R[fi * KPW + ki] -= tb[TriBlock::off_phif(P) + lq * NP + k] * Fq;
            }
          }
        }
#pragma unroll
        for (int ki = 0; ki < KPW; ki++) {
          const int k = row_k(wave, ki);
          if (k < NP && valid) {
            const long i = (long)(f * NP + k) * E + e;
            const CT r = g.R[fi * KPW + ki];
            if (residual_only) {
              a.R[i] = double(r);
            } else if (lts) {
              // span increment += w h_b k; even stages write the predictor
              // y_m + h_b k, the span's last stage writes y_m + increment
              const double v = double(r) * dtj, inc = ac[ki] + a.lts_w * v;
              if (!(a.stage & LTS_ODD))
                ((OT*)a.Uout)[i] = OT(un[ki] + v);
              if (a.stage & LTS_END)
                ((OT*)a.Uout)[i] = OT(un[ki] + inc);
              else
                a.acc[i] = inc;
            } else {
              double accv = ac[ki];
              const double o = stage_value_hd(sc, double(r) * dtj, un[ki], us[ki], &accv, g.rsum);
              if (sc >= 3 * 4 && sc < 3 * 4 + 3) a.acc[i] = accv;
              ((OT*)a.Uout)[i] = OT(o);
              if constexpr (sizeof(CT) == 4 && sizeof(OT) == 8) {
                if (a.Uout32) a.Uout32[i] = float(o);
              }
            }
          }
        }
      }
    }
  }
};

}  // namespace dgp
