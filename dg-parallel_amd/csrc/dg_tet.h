// Tiled residual + RK stage kernel for tetrahedra, Maxwell's equations in 3D
// (eps = mu = 1; MaxwellLawD).  The reference has Maxwell in 2D only
// (MaxwellLaw.cc:21-200); the element and face terms follow its DG cells
// (DGCell.cc:184-249 volume, 639-732 faces) with the tetrahedron rules the
// reference selects (GaussQuadratureTet.cc:349-380 volume, triangle rule of
// degree 2p+1 on faces).
//
// 64 elements per 256-thread workgroup, lane = element, waves share rows (field,
// mode) or face points.  Phases:
//   A   load the tile's coefficients (rows of 64 consecutive doubles), |J| J^-1
//       and the face links into LDS;
//   B   volume term in operator form: the Maxwell flux is linear, F_d(u) = A_d u,
//       so int F . grad phi_k = sum_r sum_d mg[r][d] sum_j K_r[k][j] (A_d U_j) with
//       K_r the reference stiffness operators (exact for the degree-2p rule);
//   per face l (one at a time, so LDS holds one face's traces):
//   C1  neighbour traces at the 13 face points, in the neighbour's basis through
//       the trace table of its local face and the face's vertex permutation;
//   C2  own trace, numerical flux (PEC faces: ghost E -> -E, H -> H), times the
//       face area and weight, written over the neighbour trace of that point;
//   L   lift into the register rows;
//   D   inverse mass 1/|J| and the RK stage update.
#pragma once

#include "dg_kernels.h"
#include "dg_laws.h"
#include "dg_stage_tiled.h"
#include "dg_tables.h"

namespace dgp {

struct TetArgs {
  int E, P;
  const double* tab;     // TetBlock of order P
  const double* mg;      // [9][E] |J| J^-1 (r*3+d)
  const double* detj;    // [E]
  const int32_t* link;   // [4][E]
  LawParams law;
  const double* Uin;
  const double* Un;
  double* acc;
  double* Uout;
  double* R;
  double dt;
  int scheme, stage;     // scheme MODE_RESIDUAL: write R only
  double* resid_part;    // per tile (or null)
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

template <int P>
struct TiledTet {
  using LAW = MaxwellLawD;
  static constexpr int NP = np_tet(P), NQF = nqf_tet(P), NF = 6;
  static constexpr int T = TILE_ELEMS;
  static constexpr int NROW = NF * NP;
  static constexpr int RPW = (NROW + 3) / 4;
  static constexpr int FW = NQF * NF;

  struct Smem {
    double Us[NROW][T];
    double Ms[9][T];
    int Ls[4][T];
    double W[FW][T];
    double red[TILE_THREADS];
  };
  struct Regs {
    double R[RPW];
    double rsum;
  };

  DG_HD static void phaseA(const TetArgs& a, Smem& s, int tid, int tile) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const int E = a.E, e = tile * T + t;
    const bool valid = e < E;
    for (int r = wave; r < NROW; r += 4) s.Us[r][t] = valid ? a.Uin[(long)r * E + e] : 0.0;
    for (int r = wave; r < 9; r += 4) s.Ms[r][t] = valid ? a.mg[(long)r * E + e] : (r % 4 == 0 ? 1.0 : 0.0);
    s.Ls[wave][t] = valid ? a.link[(long)wave * E + e] : -BC_PEC;
  }

  // volume term, operator form
  DG_HD static void phaseB(const TetArgs& a, Smem& s, int tid, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* K = a.tab + TetBlock::off_K(P);
    for (int j = 0; j < RPW; j++) {
      const int row = wave + 4 * j;
      double acc = 0.0;
      if (row < NROW) {
        const int f = row / NP, k = row - f * NP;
        const MaxwellJac J = maxwell_jac(f);
        for (int m = 0; m < 2; m++) {
          const int gg = J.g[m], d = J.d[m];
          double dot[3] = {0.0, 0.0, 0.0};
#pragma unroll 2
          for (int jj = 0; jj < NP; jj++) {
            const double u = s.Us[gg * NP + jj][t];
#pragma unroll
            for (int r = 0; r < 3; r++) dot[r] += K[(r * NP + k) * NP + jj] * u;
          }
          acc += J.s[m] * (s.Ms[0 * 3 + d][t] * dot[0] + s.Ms[1 * 3 + d][t] * dot[1] + s.Ms[2 * 3 + d][t] * dot[2]);
        }
      }
      g.R[j] = acc;
    }
  }

  // face l: neighbour traces into W[q*NF + f]
  DG_HD static void phaseC1(const TetArgs& a, Smem& s, int tid, int tile, int l) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const bool valid = tile * T + t < a.E;
    const int link = s.Ls[l][t];
    for (int f = wave; f < NF; f += 4) {
      double v[NP];
      const bool inner = valid && link >= 0;
      const int nb = inner ? (link >> 5) : 0;
      for (int k = 0; k < NP; k++) v[k] = inner ? a.Uin[(long)(f * NP + k) * a.E + nb] : 0.0;
      const int o = (link >> 2) & 7, ln = link & 3;
      const double* tn = a.tab + TetBlock::off_phin(P) + (inner ? (ln * 6 + o) * NQF * NP : 0);
#pragma unroll 1
      for (int q = 0; q < NQF; q++) {
        double acc = 0.0;
        for (int k = 0; k < NP; k++) acc += tn[q * NP + k] * v[k];
        s.W[q * NF + f][t] = acc;
      }
    }
  }

  // face l points: flux times area and weight, over the neighbour traces
  DG_HD static void phaseC2(const TetArgs& a, Smem& s, int tid, int l) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = a.tab;
    const int link = s.Ls[l][t];
    double Nr[3], N[3];
    tet_ref_normal(l, Nr);
    for (int d = 0; d < 3; d++) N[d] = s.Ms[0 * 3 + d][t] * Nr[0] + s.Ms[1 * 3 + d][t] * Nr[1] + s.Ms[2 * 3 + d][t] * Nr[2];
    const double area = dg_sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    const double n[3] = {N[0] / area, N[1] / area, N[2] / area};
    for (int q = wave; q < NQF; q += 4) {
      double uL[NF], uR[NF], Fh[NF];
      for (int f = 0; f < NF; f++) {
        double acc = 0.0;
        for (int k = 0; k < NP; k++) acc += tb[TetBlock::off_phif(P) + (l * NQF + q) * NP + k] * s.Us[f * NP + k][t];
        uL[f] = acc;
        uR[f] = s.W[q * NF + f][t];
      }
      if (link < 0) LAW::ghost(a.law, -link, n, uL, uR);
      LAW::riemann(a.law, n, uL, uR, Fh);
      const double sw = area * tb[TetBlock::off_wf(P) + q];
      for (int f = 0; f < NF; f++) s.W[q * NF + f][t] = Fh[f] * sw;
    }
  }

  DG_HD static void phaseLift(const TetArgs& a, Smem& s, int tid, int l, Regs& g) {
    const int t = tid & (T - 1), wave = wave_uniform(tid >> 6);
    const double* tb = a.tab + TetBlock::off_phif(P) + l * NQF * NP;
    for (int j = 0; j < RPW; j++) {
      const int row = wave + 4 * j;
      if (row < NROW) {
        const int f = row / NP, k = row - f * NP;
        double acc = g.R[j];
#pragma unroll 1
        for (int q = 0; q < NQF; q++) acc -= tb[q * NP + k] * s.W[q * NF + f][t];
        g.R[j] = acc;
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
    for (int j = 0; j < RPW; j++) {
      const int row = wave + 4 * j;
      if (row >= NROW) continue;
      const long i = (long)row * E + e;
      if (a.scheme == MODE_RESIDUAL) {
        a.R[i] = g.R[j];
        continue;
      }
      double accv = reads_acc ? a.acc[i] : 0.0;
      const double o = stage_value_hd(sc, g.R[j] * dtj, a.Un[i], reads_input ? a.Uin[i] : 0.0, &accv, g.rsum);
      if (sc >= 3 * 4 && sc < 3 * 4 + 3) a.acc[i] = accv;
      a.Uout[i] = o;
    }
  }
};

int launch_tet_stage(const TetArgs& a, hipStream_t s);
int tet_tiles(int E);

}  // namespace dgp
