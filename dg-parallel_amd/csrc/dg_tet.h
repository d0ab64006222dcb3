// Tiled residual + RK stage kernel for tetrahedra, Maxwell's equations in 3D
// (eps = mu = 1; MaxwellLawD).  The reference has Maxwell in 2D only
// (MaxwellLaw.cc:21-200); the element and face terms follow its DG cells
// (DGCell.cc:184-249 volume, 639-732 faces) with the tetrahedron rules the
// reference selects (GaussQuadratureTet.cc:349-380 volume, triangle rule of
// degree 2p+1 on faces).
//
// 64 elements per 256-thread workgroup, lane = element, waves share rows (field,
// mode) or face points.  Phases:
//   A   load the tile's coefficients (rows of 64 consecutive doubles) and the
//       face links into LDS;
//   B   volume term in operator form: the Maxwell flux is linear, F_d(u) = A_d u,
//       so int F . grad phi_k = sum_r sum_d mg[r][d] sum_j K_r[k][j] (A_d U_j) with
//       K_r the reference stiffness operators (exact for the degree-2p rule) and
//       mg = |J| J^-1 read per lane;
//   per face l (one at a time; the two trace arrays of a face reuse the
//   coefficient rows once B is done, which keeps LDS at 80 KB: two workgroups
//   per CU):
//   C1  own and neighbour traces at the face points (12 tasks, field x side,
//       dealt over the waves), each from global coefficients (own rows are coalesced,
//       neighbour rows gathered) through the trace table of the local face, and
//       for the neighbour of its local face and the face's vertex permutation;
//   C2  numerical flux (PEC faces: ghost E -> -E, H -> H), times the face area
//       and weight, written over the own trace of that point;
//   L   lift into the register rows;
//   D   inverse mass 1/|J| and the RK stage update.
#pragma once

#include <type_traits>

#include "dg_kernels.h"
#include "dg_laws.h"
#include "dg_stage_tiled.h"
#include "dg_tables.h"

namespace dgp {

struct TetArgs {
  int E, P;
  const double* tab;     // TetBlock of order P
  const float* tabf;     // its FP32 copy (the FP32 form)
  const double* mg;      // [9][E] |J| J^-1 (r*3+d)
  const double* detj;    // [E]
  const int32_t* link;   // [4][E]
  LawParams law;
  const void* Uin;       // stage input: FP64, or FP32 in the FP32 form's stages 2..4
  const double* Un;
  double* acc;
  void* Uout;            // stage output: FP64, or FP32 in the FP32 form's stages 1..3
  double* R;
  double dt;
  int scheme, stage;     // scheme MODE_RESIDUAL: write R only
  double* resid_part;    // per tile (or null)
  // FP32 form, the last stage: the FP64 output rounded to FP32 as well, the
  // next step's first-stage input (null: not written)
  float* Uout32;
};

// d F[f][d] / d U[g] of MaxwellLawD::flux as (d, g, sign) pairs per field
struct MaxwellJac {
  int d[2], g[2];
  double s[2];
};
DG_HD MaxwellJac maxwell_jac(int f) {
  switch (f) {
    case 0: return {{1, 2}, {5, 4}, {-1.0, 1.0}};
    case 1: return {{0, 2}, {5, 3}, {1.0, -1.0}};
    case 2: return {{0, 1}, {4, 3}, {-1.0, 1.0}};
    case 3: return {{1, 2}, {2, 1}, {1.0, -1.0}};
    case 4: return {{0, 2}, {2, 0}, {-1.0, 1.0}};
    default: return {{0, 1}, {1, 0}, {1.0, -1.0}};
  }
}

// area-weighted outward normals of the reference tet's faces {0,1,2}, {0,1,3},
// {1,2,3}, {0,2,3}
DG_HD void tet_ref_normal(int l, double N[3]) {
  N[0] = l == 2 ? 0.5 : l == 3 ? -0.5 : 0.0;
  N[1] = l == 2 ? 0.5 : l == 1 ? -0.5 : 0.0;
  N[2] = l == 2 ? 0.5 : l == 0 ? -0.5 : 0.0;
}

// NW waves share a tile of 64 elements (4: 256 threads, 8: 512, 16: 1024, with
// fewer rows per thread).  ALIAS: the face traces reuse the coefficient rows in
// LDS (80 KB at P3 in FP64: two workgroups per CU; own traces are then taken
// from the coefficients in global memory, one pass per face), else the
// coefficient rows stay and own traces come from LDS (101 KB at P3 in FP64: one
// workgroup per CU; 50 KB in FP32: three).
// CT: arithmetic type (the FP32 form: float, with the FP32 tables); IT / OT:
// stage input / output types; the step-start state, the accumulator, dt/|J| and
// the stage update stay FP64.
template <int P, int NW = 4, bool ALIAS = true, typename CT = double, typename IT = double, typename OT = double>
struct TiledTet {
  using LAW = MaxwellLawD;
  static constexpr int NP = np_tet(P), NQF = nqf_tet(P), NF = 6;
  static constexpr int T = TILE_ELEMS;
  static constexpr int THREADS = NW * T;
  static constexpr int NROW = NF * NP;
  static constexpr int FW = NQF * NF;

  struct Traces {
    CT WL[FW][T];  // own trace, then flux times area and weight
    CT WR[FW][T];  // neighbour trace
  };
  struct SmemAlias {
    union {
      CT Us[NROW][T];  // phases A-B
      Traces W;        // phases C1-L
      double red[THREADS];
    };
    int Ls[4][T];
  };
  struct SmemKeep {
    union {
      struct {
        CT Us[NROW][T];
        CT WR[FW][T];  // neighbour trace, then flux times area and weight
      };
      double red[THREADS];  // after the stage update
    };
    int Ls[4][T];
  };
  static_assert(ALIAS || sizeof(CT) * (NROW + FW) * T >= sizeof(double) * THREADS, "reduction buffer");
  using Smem = std::conditional_t<ALIAS, SmemAlias, SmemKeep>;
  // the rows that hold the neighbour trace and the flux of face point row r
  DG_HD static CT* nb_row(Smem& s, int r) {
    if constexpr (ALIAS) return s.W.WR[r]; else return s.WR[r];
  }
  DG_HD static CT* flux_row(Smem& s, int r) {
    if constexpr (ALIAS) return s.W.WL[r]; else return s.WR[r];
  }
  DG_HD static const CT* table(const TetArgs& a) {
    if constexpr (sizeof(CT) == 8) return (const CT*)(const void*)a.tab; else return (const CT*)(const void*)a.tabf;
  }
  struct Regs {
    CT R[NF * ((NP + NW - 1) / NW)];
    double rsum;
  };

  DG_HD static void phaseA(const TetArgs& a, Smem& s, int tid, int tile) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, e = tile * T + t;
    const bool valid = e < E;
    const int es = valid ? e : 0;  // padding lanes load element 0 (in bounds) and discard it
    // every row load is issued, unconditionally, before the LDS writes use any
    constexpr int RA = (NROW + NW - 1) / NW;
    const IT* Uin = (const IT*)a.Uin;
    CT u[RA];
#pragma unroll
    for (int j = 0; j < RA; j++) {
      const int r = wave + NW * j;
      u[j] = r < NROW ? CT(Uin[(long)r * E + es]) : CT(0);
    }
    const int lk = a.link[(long)(wave & 3) * E + es];
#pragma unroll
    for (int j = 0; j < RA; j++) {
      const int r = wave + NW * j;
      if (r < NROW) s.Us[r][t] = valid ? u[j] : CT(0);
    }
    if (wave < 4) s.Ls[wave][t] = valid ? lk : -BC_PEC;
  }

  // |J| J^-1 of the lane's element (the identity past the end of the mesh)
  DG_HD static void load_mg(const TetArgs& a, int tile, int t, CT m[3][3]) {
    const int E = a.E, e = tile * T + t;
    const bool valid = e < E;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int d = 0; d < 3; d++) m[r][d] = valid ? CT(a.mg[(long)(r * 3 + d) * E + e]) : CT(r == d ? 1 : 0);
  }

  // Rows are owned by k-slot: wave w holds the modes k = w + NW ks (ks < KS) of
  // every field, so row (f, k) is register j = f*KS + ks and the table rows a
  // wave reads depend on ks only (wave-uniform, scalar loads).
  static constexpr int KS = (NP + NW - 1) / NW;
  DG_HD static int kmode(int wave, int ks) { return wave + NW * ks; }

  // volume term, operator form: for each of the wave's modes k, the 3 x 6 products
  // (K_r U_g)[k] once, then every field's row from its two (d, g) pairs
  DG_HD static void phaseB(const TetArgs& a, Smem& s, int tid, int tile, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const CT* K = table(a) + TetBlock::off_K(P);
    CT m[3][3];
    load_mg(a, tile, t, m);
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int k = kmode(wave, ks);
      if (k >= NP) {
#pragma unroll
        for (int f = 0; f < NF; f++) g.R[f * KS + ks] = CT(0);
        continue;
      }
      CT dot[NF][3];
#pragma unroll
      for (int gg = 0; gg < NF; gg++)
#pragma unroll
        for (int r = 0; r < 3; r++) dot[gg][r] = CT(0);
#pragma unroll 4
      for (int jj = 0; jj < NP; jj++) {
        const CT k0 = K[(0 * NP + k) * NP + jj], k1 = K[(1 * NP + k) * NP + jj], k2 = K[(2 * NP + k) * NP + jj];
#pragma unroll
        for (int gg = 0; gg < NF; gg++) {
          const CT u = s.Us[gg * NP + jj][t];
          dot[gg][0] += k0 * u;
          dot[gg][1] += k1 * u;
          dot[gg][2] += k2 * u;
        }
      }
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const MaxwellJac J = maxwell_jac(f);
        CT acc = CT(0);
#pragma unroll
        for (int mm = 0; mm < 2; mm++) {
          const int gg = J.g[mm], d = J.d[mm];
          acc += CT(J.s[mm]) * (m[0][d] * dot[gg][0] + m[1][d] * dot[gg][1] + m[2][d] * dot[gg][2]);
        }
        g.R[f * KS + ks] = acc;
      }
    }
  }

  // face l: own traces into WL[q*NF + f], neighbour traces into WR[q*NF + f]
  DG_HD static void phaseC1(const TetArgs& a, Smem& s, int tid, int tile, int l) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, e = tile * T + t;
    const bool valid = e < E;
    const int es = valid ? e : 0;  // padding lanes read element 0 (in bounds) and discard it
    const int link = s.Ls[l][t];
    const bool inner = valid && link >= 0;
    const int nb = inner ? (link >> 5) : es;
    // ALIAS: 12 tasks, own then neighbour traces; else the 6 neighbour traces.
    // Software-pipelined: the next task's coefficient rows are in flight while
    // the current task's traces are formed.
    constexpr int T0 = ALIAS ? 0 : NF;
    constexpr int NT = (2 * NF - T0 + NW - 1) / NW;  // tasks of a wave, at most
    const IT* Uin = (const IT*)a.Uin;
    const CT* tab = table(a);
    auto load = [&](int j, CT* v) {
      const int task = T0 + wave + NW * j;
      if (task >= 2 * NF) return;
      const int f = task < NF ? task : task - NF;
      const int src = task < NF ? es : nb;
#pragma unroll
      for (int k = 0; k < NP; k++) v[k] = CT(Uin[(long)(f * NP + k) * E + src]);
    };
    CT vb[2][NP];
    load(0, vb[0]);
#pragma unroll
    for (int j = 0; j < NT; j++) {
      if (j + 1 < NT) load(j + 1, vb[(j + 1) & 1]);
      const int task = T0 + wave + NW * j;
      if (task >= 2 * NF) break;
      CT* v = vb[j & 1];
      if (task < NF) {
        const int f = task;
        const CT* tn = tab + TetBlock::off_phif(P) + l * NQF * NP;
        for (int q = 0; q < NQF; q++) {
          CT acc = CT(0);
#pragma unroll
          for (int k = 0; k < NP; k++) acc += tn[q * NP + k] * v[k];
          flux_row(s, q * NF + f)[t] = acc;
        }
        continue;
      }
      const int f = task - NF;
#pragma unroll
      for (int k = 0; k < NP; k++) v[k] = inner ? v[k] : CT(0);
      const int key = inner ? ((link >> 2) & 7) * 4 + (link & 3) : 0;  // perm * 4 + neighbour face
      const int key0 = wave_uniform(key);
      if (wave_all(key == key0)) {
        // the usual case (elements are grouped by link signature): one table
        // for the wave, read through the scalar cache
        const CT* tn = tab + TetBlock::off_phin(P) + ((key0 & 3) * 6 + (key0 >> 2)) * NQF * NP;
        for (int q = 0; q < NQF; q++) {
          CT acc = CT(0);
#pragma unroll
          for (int k = 0; k < NP; k++) acc += tn[q * NP + k] * v[k];
          nb_row(s, q * NF + f)[t] = acc;
        }
      } else {
        const CT* tn = tab + TetBlock::off_phin(P) + ((key & 3) * 6 + (key >> 2)) * NQF * NP;
#pragma unroll 1
        for (int q = 0; q < NQF; q++) {
          CT acc = CT(0);
          for (int k = 0; k < NP; k++) acc += tn[q * NP + k] * v[k];
          nb_row(s, q * NF + f)[t] = acc;
        }
      }
    }
  }

  // face l points: flux times area and weight, over the neighbour traces
  DG_HD static void phaseC2(const TetArgs& a, Smem& s, int tid, int tile, int l) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const CT* tb = table(a);
    const int link = s.Ls[l][t];
    CT m[3][3], N[3];
    double Nr[3];
    load_mg(a, tile, t, m);
    tet_ref_normal(l, Nr);
    for (int d = 0; d < 3; d++) N[d] = m[0][d] * CT(Nr[0]) + m[1][d] * CT(Nr[1]) + m[2][d] * CT(Nr[2]);
    const CT area = dg_sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    const CT n[3] = {N[0] / area, N[1] / area, N[2] / area};
    for (int q = wave; q < NQF; q += NW) {
      CT uL[NF], uR[NF], Fh[NF];
      if constexpr (ALIAS) {
        for (int f = 0; f < NF; f++) uL[f] = s.W.WL[q * NF + f][t];
      } else {
        const CT* to = tb + TetBlock::off_phif(P) + (l * NQF + q) * NP;
        for (int f = 0; f < NF; f++) {
          CT acc = CT(0);
#pragma unroll
          for (int k = 0; k < NP; k++) acc += to[k] * s.Us[f * NP + k][t];
          uL[f] = acc;
        }
      }
      for (int f = 0; f < NF; f++) uR[f] = nb_row(s, q * NF + f)[t];
      if (link < 0) LAW::ghost(a.law, -link, n, uL, uR);
      LAW::riemann(a.law, n, uL, uR, Fh);
      const CT sw = area * tb[TetBlock::off_wf(P) + q];
      for (int f = 0; f < NF; f++) flux_row(s, q * NF + f)[t] = Fh[f] * sw;
    }
  }

  DG_HD static void phaseLift(const TetArgs& a, Smem& s, int tid, int l, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const CT* tb = table(a) + TetBlock::off_phif(P) + l * NQF * NP;
    for (int q = 0; q < NQF; q++) {
      CT Fq[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) Fq[f] = flux_row(s, q * NF + f)[t];
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int k = kmode(wave, ks);
        if (k < NP) {
          const CT phi = tb[q * NP + k];
#pragma unroll
          for (int f = 0; f < NF; f++) g.R[f * KS + ks] -= phi * Fq[f];
        }
      }
    }
  }

  DG_HD static void phaseD(const TetArgs& a, int tid, int tile, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, e = tile * T + t;
    g.rsum = 0.0;
    if (e >= E) return;
    const int sc = a.scheme * 4 + a.stage;
    const bool reads_input = sc == 2 * 4 + 1 || sc == 2 * 4 + 2 || sc == 1 * 4 + 1;
    const bool reads_acc = sc == 3 * 4 + 1 || sc == 3 * 4 + 2 || sc == 3 * 4 + 3;
    const double dtj = a.dt / a.detj[e];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int k = kmode(wave, ks);
        if (k >= NP) continue;
        const long i = (long)(f * NP + k) * E + e;
        const double r = double(g.R[f * KS + ks]);
        if (a.scheme == MODE_RESIDUAL) {
          a.R[i] = r;
          continue;
        }
        double accv = reads_acc ? a.acc[i] : 0.0;
        const double o = stage_value_hd(sc, r * dtj, a.Un[i], reads_input ? double(((const IT*)a.Uin)[i]) : 0.0,
                                        &accv, g.rsum);
        if (sc >= 3 * 4 && sc < 3 * 4 + 3) a.acc[i] = accv;
        ((OT*)a.Uout)[i] = OT(o);
        if constexpr (sizeof(CT) == 4 && sizeof(OT) == 8) {
          if (a.Uout32) a.Uout32[i] = float(o);
        }
      }
  }
};

// shape: 0 for 4 waves per tile, 1 for 8 (traces over the coefficient rows);
// 2 for 8, 3 for 16, 4 for 4 (coefficient rows kept: no global re-reads of
// the own rows; FP32 LDS 49 KB at P3, three workgroups per CU).  prec: 0 FP64;
// 1..3 the FP32 form (1: FP64 input, FP32 output; 2: FP32 both; 3: FP32 input,
// FP64 output).
constexpr int TET_SHAPES = 5;
int launch_tet_stage(const TetArgs& a, int shape, int prec, hipStream_t s);
int launch_tet_stage32(const TetArgs& a, int shape, int prec, hipStream_t s);
int tet_tiles(int E);

}  // namespace dgp
