// Tiled mixed-order (P1-P4) triangle stage.  Work items are tiles of 64
// elements of one order bucket, run either by one persistent launch over every
// bucket (dg_mixed.hip walks the chunk list) or by one launch per bucket (LDS
// rows and registers of that order; PF: C1's rows and the stage update's
// operands loaded early).  Per tile of order P, 256 threads, lane = element,
// the four waves share rows (field, mode) or points as the uniform kernel does
// (dg_stage_tiled.h):
//   A   the tile's coefficients into LDS (rows of 64 consecutive doubles of
//       bucket P), |J| J^-1, face normals and lengths, links, and per face the
//       rule size nq = max(P, P_nb) + 1 (boundary faces: P + 1; DGCell.cc:408-423);
//   B/D1 volume flux at the NQ points, MIX_QC points per pass, contracted with
//       the basis gradients into the register rows (DGCell.cc:184-249);
//   C1  neighbour traces at the nq points of each face in the neighbour's own
//       basis (its order, local face and orientation vary per lane: the trace
//       rows are read per lane from the tables staged in LDS, every lane over
//       all points and rows at in-bounds places, without branches);
//   C2  own traces and the numerical flux times length and weight, written over
//       the neighbour trace of that point (DGCell.cc:639-732);
//   D2  lift into the rows and the RK stage update.
// C2's own traces and D2's lift read the order-P trace table of rule size nq in
// wave-uniform passes over the sizes present in the wave (scalar loads, lanes of
// other sizes idle; D2's passes with their points unrolled), since nq takes at
// most MAXP_TRI + 1 - P values.
// Per-lane choices among the buckets (neighbour order) go through select chains,
// never through a per-lane index into the argument block.
#pragma once

#include "dg_mixed.h"

namespace dgp {

constexpr int MIX_TILE = 64;             // elements per work item
constexpr int MIX_NQF = MAXP_TRI + 1;    // the largest face rule
constexpr int MIX_QC = 8;                // volume points per pass

DG_HD bool wave_any(bool pred) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __any(pred);
#else
  return pred;
#endif
}

// PU: the largest order of the tiles that share it (MAXP_TRI for the launch
// over every bucket, P for a launch over bucket P alone)
template <class LAW, int PU = MAXP_TRI>
struct MixSmem {
  static constexpr int NF = LAW::NF;
  static constexpr int UROWS = NF * np_tri(PU);
  static constexpr int GW = MIX_QC * NF * 2;     // volume-flux rows of one pass
  static constexpr int FW = 3 * MIX_NQF * NF;    // face trace / flux rows
  static constexpr int WROWS = GW > FW ? GW : FW;
  static_assert(WROWS * MIX_TILE >= TILE_THREADS, "reduction buffer");
  double Us[UROWS][MIX_TILE];
  union {
    double W[WROWS][MIX_TILE];
    double red[TILE_THREADS];  // the workgroup's residual-norm reduction, after the last tile
  };
  double Ms[4][MIX_TILE];
  double Ns[3][3][MIX_TILE];  // per face: unit outward normal (2) and length
  int Ls[3][MIX_TILE];
  int Nq[3][MIX_TILE];
  double Tn[MIX_NTAB];  // neighbour trace tables, staged once per workgroup
  double Gw[GLW_OFF(MIX_NQF + 1)];  // Gauss-Legendre half weights of every face rule size
};

// bucket quantities of a per-lane order p (1..4) by select chains
DG_HD long sel_boff(const MixedArgs& a, int p) {
  return p == 1 ? a.boff[1] : p == 2 ? a.boff[2] : p == 3 ? a.boff[3] : a.boff[4];
}
DG_HD int sel_bstart(const MixedArgs& a, int p) {  // p in 1..5 (5: the end of bucket 4)
  return p == 1 ? a.bstart[1] : p == 2 ? a.bstart[2] : p == 3 ? a.bstart[3] : p == 4 ? a.bstart[4] : a.bstart[5];
}
DG_HD int sel_moff(const MixedArgs& a, int p) {
  return p == 1 ? a.moff[1] : p == 2 ? a.moff[2] : p == 3 ? a.moff[3] : a.moff[4];
}
// offset of the trace table [3][nq][Np(p)] inside the block of order p
DG_HD int trace_off(int p, int nq) {
  return nq == p + 1 ? TriBlock::off_phif(p) : TriBlock::off_phifx(p, nq);
}

// PF: C1's neighbour rows and the stage update's operands are loaded early (in
// phase A and before the lift) at the cost of registers: the launches over one
// bucket (PF = true) have them to spare, the persistent launch (every order's
// code in one kernel) keeps its register budget
template <int P, class LAW, int PU = MAXP_TRI, bool PF = false>
struct TiledMixed {
  static constexpr int NP = np_tri(P), NQ = nqv_tri(P), NF = LAW::NF;
  static constexpr int T = MIX_TILE;
  static constexpr int NROW = NF * NP;
  static_assert(P <= PU, "the LDS rows hold the tile's order");
  using Smem = MixSmem<LAW, PU>;
  // output rows of a wave as in the uniform kernel: whole fields for Euler
  // (f = wave), modes k = wave, wave+4, .. of the single scalar field
  static constexpr bool BY_FIELD = NF >= 4;
  static constexpr int FPW = BY_FIELD ? (NF + 3) / 4 : NF;
  static constexpr int KPW = BY_FIELD ? NP : (NP + 3) / 4;
  static constexpr int QPP = MIX_QC / 4;             // volume points per wave per pass
  static constexpr int CPW = (3 * NF + 3) / 4;       // neighbour-trace tasks per wave
  static constexpr int FPPW = (3 * MIX_NQF + 3) / 4; // face-point items per wave
  // the output rows, and C1's neighbour rows and links, loaded in phase A so
  // their latency hides behind B and D1 (rows past the neighbour's order are
  // loaded at an in-bounds index and discarded)
  struct Regs {
    double R[FPW * KPW];
    double rsum;
    double nv[PF ? CPW : 1][MIX_MAXNP];
    int lk[CPW];
  };

  DG_HD static int row_f(int wave, int fi) { return BY_FIELD ? wave + 4 * fi : fi; }
  DG_HD static int row_k(int wave, int ki) { return BY_FIELD ? ki : wave + 4 * ki; }
  // coefficient row r of internal element i of bucket P
  DG_HD static long idx(const MixedArgs& a, int i, int r) {
    return a.boff[P] + (long)r * (a.bstart[P + 1] - a.bstart[P]) + (i - a.bstart[P]);
  }
  DG_HD static const double* block(const MixedArgs& a) { return a.tab + a.moff[P]; }

  // the neighbour trace tables into LDS (once per workgroup, before any tile)
  DG_HD static void stage_tables(const MixedArgs& a, Smem& s, int tid) {
    for (int j = tid; j < MIX_NTAB; j += TILE_THREADS) s.Tn[j] = a.ntab[j];
    if (tid < GLW_OFF(MIX_NQF + 1)) s.Gw[tid] = a.glw[tid];
  }
  // the coefficients of field f of the neighbour across link (rows of bucket
  // pn are stride doubles apart), every load issued at an in-bounds index
  // before any is used: boundary faces read the own element is, rows past the
  // neighbour's order repeat its first row (both discarded by C1)
  DG_HD static void load_nb(const MixedArgs& a, int link, int is, int f, double* v) {
    const int nb = link >= 0 ? link >> 3 : is;
    const int pn = mixed_order_of(a, nb), npn = (pn + 1) * (pn + 2) / 2;
    const int b0 = sel_bstart(a, pn), b1 = sel_bstart(a, pn + 1);
    const long stride = b1 - b0;
    const long base = sel_boff(a, pn) + (long)f * npn * stride + (nb - b0);
#pragma unroll
    for (int k = 0; k < MIX_MAXNP; k++) v[k] = a.Uin[base + (long)(k < npn ? k : 0) * stride];
  }

  // ---- A ------------------------------------------------------------------------
  DG_HD static void phaseA(const MixedArgs& a, Smem& s, int tid, int i0, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, i = i0 + t;
    const bool valid = i < a.bstart[P + 1];
    const int is = valid ? i : a.bstart[P];  // padding lanes load the bucket's first element and discard it
    // every load is issued, unconditionally, before any is used: the links
    // first (C1's addresses need them), the rows, |J| J^-1, then C1's rows
    int lk3[3];
#pragma unroll
    for (int l = 0; l < 3; l++) lk3[l] = valid ? a.el.link[(long)l * E + is] : -BC_FARFIELD;
    constexpr int RA = (NROW + 3) / 4;
    double u[RA];
#pragma unroll
    for (int j = 0; j < RA; j++) {
      const int r = wave + 4 * j;
      u[j] = r < NROW ? a.Uin[idx(a, is, r)] : 0.0;
    }
    double m[4];
#pragma unroll
    for (int r = 0; r < 4; r++) m[r] = a.el.mg[(long)r * E + is];
#pragma unroll
    for (int j = 0; j < CPW; j++) {
      const int it = wave + 4 * j;
      const int l = it < 3 * NF ? it / NF : 0, f = it < 3 * NF ? it - l * NF : 0;
      const int link = lk3[l];
      g.lk[j] = link;
      if constexpr (PF) load_nb(a, link, is, f, g.nv[j]);
    }
#pragma unroll
    for (int j = 0; j < RA; j++) {
      const int r = wave + 4 * j;
      if (r < NROW) s.Us[r][t] = valid ? u[j] : (r == 0 ? 1.0 : 0.0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (!valid) m[r] = r == 0 || r == 3 ? 1.0 : 0.0;
    s.Ms[wave][t] = m[wave];
    if (wave < 3) {
      const int l = wave;
      const int link = lk3[l];
      s.Ls[l][t] = link;
      double mg[2][2];
      for (int r = 0; r < 2; r++)
        for (int d = 0; d < 2; d++) mg[r][d] = m[r * 2 + d];
      double n[2], len;
      edge_normal(mg, l, n, len);
      s.Ns[l][0][t] = n[0];
      s.Ns[l][1][t] = n[1];
      s.Ns[l][2][t] = len;
      const int pn = link >= 0 ? mixed_order_of(a, link >> 3) : P;
      s.Nq[l][t] = (pn > P ? pn : P) + 1;
    }
#pragma unroll
    for (int j = 0; j < FPW * KPW; j++) g.R[j] = 0.0;
  }

  // ---- B: volume flux at points [q0, q0 + MIX_QC) ---------------------------------
  DG_HD static void phaseB(const MixedArgs& a, Smem& s, int tid, int q0) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = block(a);
#pragma unroll 1
    for (int j = 0; j < QPP; j++) {
      const int qi = wave + 4 * j, q = q0 + qi;
      if (q < NQ) {
        double u[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          double acc = 0.0;
#pragma unroll
          for (int k = 0; k < NP; k++) acc += tb[TriBlock::off_phi(P) + q * NP + k] * s.Us[f * NP + k][t];
          u[f] = acc;
        }
        double F[NF][2];
        LAW::flux(a.law, u, F);
        const double w = tb[TriBlock::off_wv(P) + q];
        const double m00 = s.Ms[0][t], m01 = s.Ms[1][t], m10 = s.Ms[2][t], m11 = s.Ms[3][t];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          s.W[(qi * NF + f) * 2 + 0][t] = w * (m00 * F[f][0] + m01 * F[f][1]);
          s.W[(qi * NF + f) * 2 + 1][t] = w * (m10 * F[f][0] + m11 * F[f][1]);
        }
      }
    }
  }

  // ---- D1: contraction of one pass with the basis gradients -------------------
  DG_HD static void phaseD1(const MixedArgs& a, Smem& s, int tid, int q0, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = block(a);
    const int qn = NQ - q0 < MIX_QC ? NQ - q0 : MIX_QC;
#pragma unroll
    for (int fi = 0; fi < FPW; fi++) {
      const int f = row_f(wave, fi);
      if (f < NF) {
        for (int qi = 0; qi < qn; qi++) {
          const int q = q0 + qi;
          const double G0 = s.W[(qi * NF + f) * 2 + 0][t];
          const double G1 = s.W[(qi * NF + f) * 2 + 1][t];
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

  // ---- C1: neighbour traces at the face points ------------------------------------
  DG_HD static void phaseC1(const MixedArgs& a, Smem& s, int tid, int i0, const Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int i = i0 + t;
    const int is = i < a.bstart[P + 1] ? i : a.bstart[P];
#pragma unroll
    for (int j = 0; j < CPW; j++) {
      const int it = wave + 4 * j;
      if (it >= 3 * NF) break;
      const int l = it / NF, f = it - l * NF;
      const int link = g.lk[j], nq = s.Nq[l][t];
      double vl[MIX_MAXNP];
      if constexpr (!PF) load_nb(a, link, is, f, vl);
      const double* v = PF ? g.nv[PF ? j : 0] : vl;
      // branch-free over the lanes' neighbour orders, rule sizes, faces and
      // orientations: every lane reads MIX_NQF x MIX_MAXNP table entries at
      // in-bounds places (points past nq repeat the last, rows past the
      // neighbour's order its first) and the products it does not need are
      // zero, so the LDS reads issue back to back
      const bool has = link >= 0;
      const int nb = has ? link >> 3 : is, lb = has ? link & 3 : 0;
      const bool same = (link >> 2) & 1;
      const int pn = mixed_order_of(a, nb), npn = (pn + 1) * (pn + 2) / 2;
      const double* tn = s.Tn + ntab_off(pn, nq) + lb * nq * npn;
      double vz[MIX_MAXNP];
#pragma unroll
      for (int k = 0; k < MIX_MAXNP; k++) vz[k] = has && k < npn ? v[k] : 0.0;
#pragma unroll
      for (int q = 0; q < MIX_NQF; q++) {
        const int qv = q < nq ? q : nq - 1;
        const int qq = same ? qv : nq - 1 - qv;
        const double* row = tn + qq * npn;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < MIX_MAXNP; k++) acc += row[k < npn ? k : 0] * vz[k];
        s.W[(l * MIX_NQF + q) * NF + f][t] = q < nq ? acc : 0.0;
      }
    }
  }

  // ---- C2: own traces and the numerical flux ------------------------------------
  DG_HD static void phaseC2(const MixedArgs& a, Smem& s, int tid) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = block(a);
#pragma unroll 1
    for (int j = 0; j < FPPW; j++) {
      const int it = wave + 4 * j;
      if (it >= 3 * MIX_NQF) break;
      const int l = it / MIX_NQF, q = it - l * MIX_NQF;
      const int nq = s.Nq[l][t];
      const bool active = q < nq;
      if (!wave_any(active)) continue;
      // own trace at point q of the rule of size nq: one pass per size present
      double uL[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) uL[f] = 0.0;
#pragma unroll 1
      for (int nv = (q + 1 > P + 1 ? q + 1 : P + 1); nv <= MIX_NQF; nv++) {
        const bool mine = active && nq == nv;
        if (!wave_any(mine)) continue;
        const double* to = tb + trace_off(P, nv) + (l * nv + q) * NP;
        if (mine) {
#pragma unroll
          for (int f = 0; f < NF; f++) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < NP; k++) acc += to[k] * s.Us[f * NP + k][t];
            uL[f] = acc;
          }
        }
      }
      if (active) {
        const int link = s.Ls[l][t];
        const double n[2] = {s.Ns[l][0][t], s.Ns[l][1][t]};
        double uR[NF], Fh[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) uR[f] = s.W[(l * MIX_NQF + q) * NF + f][t];
        face_flux<LAW>(a.law, link, n, uL, uR, Fh);
        const double sw = s.Ns[l][2][t] * s.Gw[GLW_OFF(nq) + q];
#pragma unroll
        for (int f = 0; f < NF; f++) s.W[(l * MIX_NQF + q) * NF + f][t] = Fh[f] * sw;
      }
    }
  }

  // the lift of face l's fluxes at the NV points of the rule of that size, on
  // the lanes whose face l has that size
  template <int NV>
  DG_HD static void lift_nv(const MixedArgs& a, const Smem& s, int t, int wave, int l, int f, int fi, int nq,
                            Regs& g) {
    static_assert(NV <= MIX_NQF, "face rule size");
    if constexpr (NV >= P + 1) {
      const bool mine = nq == NV;
      if (!wave_any(mine)) return;
      const double* to = block(a) + trace_off(P, NV) + l * NV * NP;
      double Fq[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) Fq[q] = mine ? s.W[(l * MIX_NQF + q) * NF + f][t] : 0.0;
#pragma unroll
      for (int q = 0; q < NV; q++)
#pragma unroll
        for (int ki = 0; ki < KPW; ki++) {
          const int k = row_k(wave, ki);
          if (k < NP) g.R[fi * KPW + ki] -= to[q * NP + k] * Fq[q];
        }
    }
  }

  // ---- D2: lift and the stage update -------------------------------------------
  DG_HD static void phaseD2(const MixedArgs& a, Smem& s, int tid, int i0, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int i = i0 + t;
    const bool valid = i < a.bstart[P + 1];
    const int is = valid ? i : a.bstart[P];
    const double* tb = block(a);
    const int sc = a.scheme * 4 + a.stage;
    const bool residual_only = a.scheme == MODE_RESIDUAL;
    const bool reads_input = sc == 2 * 4 + 1 || sc == 2 * 4 + 2 || sc == 1 * 4 + 1;
    const bool reads_acc = sc == 3 * 4 + 1 || sc == 3 * 4 + 2 || sc == 3 * 4 + 3;
    const double dtj = a.dt / a.el.detj[is];
    g.rsum = 0.0;
#pragma unroll
    for (int fi = 0; fi < FPW; fi++) {
      const int f = row_f(wave, fi);
      if (f >= NF) continue;
      // PF: issue the stage update's loads first; the lift below hides them
      double un[PF ? KPW : 1], ac[PF ? KPW : 1], us[PF ? KPW : 1];
      if constexpr (PF) {
#pragma unroll
        for (int ki = 0; ki < KPW; ki++) {
          const int k = row_k(wave, ki);
          const long ix = idx(a, is, f * NP + (k < NP ? k : 0));
          un[ki] = residual_only ? 0.0 : a.Un[ix];
          ac[ki] = reads_acc ? a.acc[ix] : 0.0;
          us[ki] = reads_input ? a.Uin[ix] : 0.0;
        }
      }
#pragma unroll 1
      for (int l = 0; l < 3; l++) {
        const int nq = s.Nq[l][t];
        // one pass per rule size present in the wave, each with its points
        // unrolled (the flux reads and table rows issue together)
        switch (P + 1) {
          case 2: lift_nv<2>(a, s, t, wave, l, f, fi, nq, g); [[fallthrough]];
          case 3: lift_nv<3>(a, s, t, wave, l, f, fi, nq, g); [[fallthrough]];
          case 4: lift_nv<4>(a, s, t, wave, l, f, fi, nq, g); [[fallthrough]];
          default: lift_nv<5>(a, s, t, wave, l, f, fi, nq, g);
        }
      }
      if (!valid) continue;
#pragma unroll
      for (int ki = 0; ki < KPW; ki++) {
        const int k = row_k(wave, ki);
        if (k >= NP) continue;
        const long ix = idx(a, i, f * NP + k);
        const double r = g.R[fi * KPW + ki];
        if (residual_only) {
          a.R[ix] = r;
          continue;
        }
        const int kp = PF ? ki : 0;
        double accv, unv, usv;
        if constexpr (PF) {
          accv = ac[kp];
          unv = un[kp];
          usv = us[kp];
        } else {
          accv = reads_acc ? a.acc[ix] : 0.0;
          unv = a.Un[ix];
          usv = reads_input ? a.Uin[ix] : 0.0;
        }
        const double o = stage_value_hd(sc, r * dtj, unv, usv, &accv, g.rsum);
        if (sc >= 3 * 4 && sc < 3 * 4 + 3) a.acc[ix] = accv;
        a.Uout[ix] = o;
      }
    }
  }
};

}  // namespace dgp
