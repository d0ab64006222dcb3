// CPU emulation of the tiled triangle stage kernel, the element-per-work-item
// stage kernel and the residual-derivative kernels: runs the kernels' own phase
// and element functions (dg-parallel_amd/csrc/dg_stage_tiled.h, dg_element.h,
// dg_adjoint.h) thread by thread on the host,
// with the same host-built tables the library uploads.  Test infrastructure:
// lets the CPU test suite check the kernel arithmetic and indexing against the
// oracle without a GPU.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../dg-parallel_amd/csrc/dg_adjoint.h"
#include "../../dg-parallel_amd/csrc/dg_element.h"
#include "../../dg-parallel_amd/csrc/dg_hang.h"
#include "../../dg-parallel_amd/csrc/dg_mixed.h"
#include "../../dg-parallel_amd/csrc/dg_mixed_tiled.h"
#include "../../dg-parallel_amd/csrc/dg_tet.h"
#include "../../dg-parallel_amd/csrc/dg_setup.h"
#include "../../dg-parallel_amd/csrc/dg_stage_tiled.h"

using namespace dgp;

namespace {

template <int P, class LAW, typename CT, typename IT, typename OT, int NW>
void run_tiles(const StageArgs& a) {
  using K = TiledStage<P, LAW, CT, IT, OT, NW>;
  constexpr int TH = K::THREADS;
  std::unique_ptr<typename K::Smem> sm(new typename K::Smem());
  std::vector<typename K::Regs> g(TH);
  const int ntiles = ((a.e1 > 0 ? a.e1 : a.E) - a.e0 + TILE_ELEMS - 1) / TILE_ELEMS;
  for (int tile = 0; tile < ntiles; tile++) {
    std::memset(sm.get(), 0, sizeof(typename K::Smem));
    for (int t = 0; t < TH; t++) K::phaseA(a, *sm, t, tile, g[t]);
    for (int q0 = 0; q0 < K::NQ; q0 += K::QH) {
      for (int t = 0; t < TH; t++) K::phaseB(a, *sm, t, g[t], q0);
      for (int t = 0; t < TH; t++) K::phaseD1(a, *sm, t, g[t], q0);
    }
    for (int t = 0; t < TH; t++) K::phaseC1(a, *sm, t, tile, g[t]);
    for (int t = 0; t < TH; t++) K::phaseC2(a, *sm, t, g[t]);
    double r = 0;
    for (int t = 0; t < TH; t++) {
      K::phaseD2(a, *sm, t, tile, g[t]);
      r += g[t].rsum;
    }
    if (a.resid_part) a.resid_part[tile] = r;
  }
}

template <class LAW, typename CT, typename IT, typename OT, int NW>
int run_law(const StageArgs& a) {
  switch (a.P) {
    case 1: run_tiles<1, LAW, CT, IT, OT, NW>(a); return 0;
    case 2: run_tiles<2, LAW, CT, IT, OT, NW>(a); return 0;
    case 3: run_tiles<3, LAW, CT, IT, OT, NW>(a); return 0;
    case 4: run_tiles<4, LAW, CT, IT, OT, NW>(a); return 0;
  }
  return -1;
}

template <typename CT, typename IT, typename OT, int NW>
int run(const StageArgs& a, int law) {
  if (law == LAW_SCALAR && a.pvel) return run_law<ScalarFieldLawD, CT, IT, OT, NW>(a);
  if (law == LAW_SCALAR) return run_law<ScalarLawD, CT, IT, OT, NW>(a);
  if (law == LAW_EULER2D && a.law.char_bc) return run_law<EulerCharLawD, CT, IT, OT, NW>(a);
  if (law == LAW_EULER2D) return run_law<EulerLawD, CT, IT, OT, NW>(a);
  if (law == LAW_MAXWELL2D) return run_law<Maxwell2dLawD, CT, IT, OT, NW>(a);
  if (law == LAW_BURGERS2D) return run_law<Burgers2DLawD, CT, IT, OT, NW>(a);
  return -1;
}

template <int NW>
int run_prec(const StageArgs& a, int law, int prec) {
  switch (prec) {
    case 0: return run<double, double, double, NW>(a, law);
    case 1: return run<float, double, float, NW>(a, law);
    case 2: return run<float, float, float, NW>(a, law);
    case 3: return run<float, float, double, NW>(a, law);
  }
  return -1;
}

// variant as the device launcher's: bits 0-1 precision form, bit 2 tile shape
int run_variant(const StageArgs& a, int law, int variant) {
  switch (variant >> 2) {
    case 0: return run_prec<4>(a, law, variant & 3);
    case 1: return run_prec<8>(a, law, variant & 3);
  }
  return -1;
}

struct Emu {
  TriSetup S;
  std::vector<double> Un, acc, R, A64, B64, part;
  std::vector<float> A32, B32;
  // the hanging-face pass's tables (meshes with hanging nodes)
  std::vector<int32_t> helem, hrstart, hrec;
  std::vector<double> hph;
  StageArgs base() {
    StageArgs a{};
    a.E = S.E;
    a.P = S.P;
    a.tab = S.tab.data();
    a.tabf = S.tabf.data();
    a.el.mg = S.mg.data();
    a.el.detj = S.detj.data();
    a.el.size = S.size.data();
    a.el.link = S.link.data();
    a.law = S.lp;
    a.Un = Un.data();
    a.acc = acc.data();
    return a;
  }
  AdjArgs adj() {
    AdjArgs a{};
    a.E = S.E;
    a.P = S.P;
    a.tab = S.tab.data();
    a.el.mg = S.mg.data();
    a.el.detj = S.detj.data();
    a.el.size = S.size.data();
    a.el.link = S.link.data();
    a.law = S.lp;
    a.mode = ADJ_PLAIN;
    return a;
  }
  void to_soa(const double* U) { to_soa(U, Un.data()); }
  void to_soa(const double* U, double* dst) {
    const int BS = S.NF * S.NP;
    for (int i = 0; i < S.E; i++)
      for (int j = 0; j < BS; j++) dst[(size_t)j * S.E + i] = U[(size_t)S.perm[i] * BS + j];
  }
  void from_soa(const double* src, double* U) {
    const int BS = S.NF * S.NP;
    for (int i = 0; i < S.E; i++)
      for (int j = 0; j < BS; j++) U[(size_t)S.perm[i] * BS + j] = src[(size_t)j * S.E + i];
  }
};

template <int P, class LAW, int ND>
void run_vjp_tiles(const AdjArgs& a) {
  using K = TiledVjp<P, LAW, ND>;
  std::unique_ptr<typename K::Smem> sm(new typename K::Smem());
  std::vector<typename K::Regs> g(TILE_THREADS);
  const int ntiles = (a.E + TILE_ELEMS - 1) / TILE_ELEMS;
  for (int tile = 0; tile < ntiles; tile++) {
    std::memset(sm.get(), 0, sizeof(typename K::Smem));
    for (int t = 0; t < TILE_THREADS; t++) K::phaseA(a, *sm, t, tile);
    for (int t = 0; t < TILE_THREADS; t++) K::phaseB(a, *sm, t);
    for (int t = 0; t < TILE_THREADS; t++) K::phaseD1(a, *sm, t, g[t]);
    for (int l = 0; l < 3; l++) {
      for (int t = 0; t < TILE_THREADS; t++) K::phaseC1(a, *sm, t, tile, l);
      for (int t = 0; t < TILE_THREADS; t++) K::phaseC2(a, *sm, t, l);
      for (int t = 0; t < TILE_THREADS; t++) K::phaseLift(a, *sm, t, l, g[t]);
    }
    for (int t = 0; t < TILE_THREADS; t++) K::phaseOut(a, t, tile, g[t]);
  }
}

template <class LAW, int ND>
int run_vjp_law(const AdjArgs& a) {
  switch (a.P) {
    case 1: run_vjp_tiles<1, LAW, ND>(a); return 0;
    case 2: run_vjp_tiles<2, LAW, ND>(a); return 0;
    case 3: run_vjp_tiles<3, LAW, ND>(a); return 0;
    case 4: run_vjp_tiles<4, LAW, ND>(a); return 0;
  }
  return -1;
}

// nd: flux tangent directions per pass (0: all fields at once)
int run_vjp(const AdjArgs& a, int law, int nd = 0) {
  if (law == LAW_SCALAR) return run_vjp_law<ScalarLawD, 1>(a);
  if (law == LAW_MAXWELL2D) return run_vjp_law<Maxwell2dLawD, 1>(a);
  if (law == LAW_BURGERS2D) return run_vjp_law<Burgers2DLawD, 1>(a);
  if (law == LAW_EULER2D && a.law.char_bc) return run_vjp_law<EulerCharLawD, 1>(a);
  if (law == LAW_EULER2D) {
    if (nd == 1) return run_vjp_law<EulerLawD, 1>(a);
    if (nd == 2) return run_vjp_law<EulerLawD, 2>(a);
    return run_vjp_law<EulerLawD, 4>(a);
  }
  return -1;
}

template <class LAW>
int run_jvp_law(const AdjArgs& a) {
  for (int e = 0; e < a.E; e++) switch (a.P) {
      case 1: jvp_element<1, LAW>(a, e); break;
      case 2: jvp_element<2, LAW>(a, e); break;
      case 3: jvp_element<3, LAW>(a, e); break;
      case 4: jvp_element<4, LAW>(a, e); break;
      default: return -1;
    }
  return 0;
}

int run_jvp(const AdjArgs& a, int law) {
  if (law == LAW_SCALAR) return run_jvp_law<ScalarLawD>(a);
  if (law == LAW_EULER2D && a.law.char_bc) return run_jvp_law<EulerCharLawD>(a);
  if (law == LAW_EULER2D) return run_jvp_law<EulerLawD>(a);
  if (law == LAW_MAXWELL2D) return run_jvp_law<Maxwell2dLawD>(a);
  if (law == LAW_BURGERS2D) return run_jvp_law<Burgers2DLawD>(a);
  return -1;
}

int make(Emu& m, const dg_mesh_desc* mesh, const dg_law_desc* law, int order, const int32_t* grp = nullptr) {
  dg_options opt{};
  opt.order = order;
  std::string err;
  if (build_tri_setup(mesh, law, &opt, m.S, err, grp) != DG_OK) return -1;
  size_t n = (size_t)m.S.E * m.S.NF * m.S.NP;
  m.Un.assign(n, 0.0);
  m.acc.assign(n, 0.0);
  m.R.assign(n, 0.0);
  m.A64.assign(n, 0.0);
  m.B64.assign(n, 0.0);
  m.A32.assign(n, 0.0f);
  m.B32.assign(n, 0.0f);
  m.part.assign((m.S.E + TILE_ELEMS - 1) / TILE_ELEMS + 1, 0.0);
  if (!m.S.topo.hang.empty()) build_hang_tables(m.S, m.helem, m.hrstart, m.hrec, m.hph);
  return 0;
}

template <int P, class LAW>
void hang_p(const HangArgs& h) {
  using K = HangBlock<P, LAW>;
  typename K::Smem sm;
  for (int i = 0; i < h.n; i++) {
    for (int l = 0; l < HANG_LANES; l++) K::hangA(h, i, l, sm);
    for (int l = 0; l < HANG_LANES; l++) K::hangB(h, i, l, sm);
    for (int l = 0; l < HANG_LANES; l++) K::hangC(h, i, l, sm);
  }
}

template <class LAW>
void hang_law(const HangArgs& h) {
  switch (h.P) {
    case 1: hang_p<1, LAW>(h); break;
    case 2: hang_p<2, LAW>(h); break;
    case 3: hang_p<3, LAW>(h); break;
    case 4: hang_p<4, LAW>(h); break;
  }
}

// the library's stage launch: the stage kernel, then (FP64, meshes with hanging
// nodes) the hanging-face pass on the same arguments
int run_stage(Emu& m, const StageArgs& a, int variant) {
  if (run_variant(a, m.S.law, variant)) return -2;
  if (m.helem.empty()) return 0;
  if ((variant & 3) != 0) return -3;
  HangArgs h{};
  h.E = m.S.E;
  h.P = m.S.P;
  h.n = (int)m.helem.size();
  h.elem = m.helem.data();
  h.rstart = m.hrstart.data();
  h.rec = m.hrec.data();
  h.Uin = (const double*)a.Uin;
  h.Uout = (double*)a.Uout;
  h.acc = a.acc;
  h.R = a.R;
  h.tab = m.S.tab.data();
  h.phih = m.hph.data();
  h.el = a.el;
  h.law = m.S.lp;
  h.dt = a.dt;
  h.scheme = a.scheme;
  h.stage = a.stage;
  const int law = m.S.law;
  if (law == LAW_SCALAR) hang_law<ScalarLawD>(h);
  else if (law == LAW_EULER2D && h.law.flux == FLUX_EXACT) hang_law<EulerExactLawD>(h);
  else if (law == LAW_EULER2D) hang_law<EulerLawD>(h);
  else if (law == LAW_MAXWELL2D) hang_law<Maxwell2dLawD>(h);
  else if (law == LAW_BURGERS2D) hang_law<Burgers2DLawD>(h);
  else return -4;
  return 0;
}

}  // namespace

// form: bit 0 the FP32 form, bit 1 the 8-wave tile shape
extern "C" int emu_tri_residual(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int form,
                                const double* U, double* R) {
  const int fp32 = form & 1, shape = (form >> 1) & 1;
  Emu m;
  if (make(m, mesh, law, order)) return -1;
  m.to_soa(U);
  StageArgs a = m.base();
  a.Uin = m.Un.data();
  a.R = m.R.data();
  a.scheme = MODE_RESIDUAL;
  if (run_stage(m, a, (fp32 ? 1 : 0) | shape << 2)) return -2;
  m.from_soa(m.R.data(), R);
  return 0;
}

template <int P, class LAW>
void element_residual(const StageArgs& a) {
  double rsum = 0.0;
  for (int e = 0; e < a.E; e++) element_stage<P, LAW, double, double, double>(a, e, rsum);
}

template <class LAW>
int element_residual_law(const StageArgs& a) {
  switch (a.P) {
    case 1: element_residual<1, LAW>(a); return 0;
    case 2: element_residual<2, LAW>(a); return 0;
    case 3: element_residual<3, LAW>(a); return 0;
    case 4: element_residual<4, LAW>(a); return 0;
  }
  return -1;
}

// the FP64 residual of the element-per-work-item kernel (every law, including
// those the library runs on the tiled kernels only)
extern "C" int emu_tri_residual_element(const dg_mesh_desc* mesh, const dg_law_desc* law, int order,
                                        const double* U, double* R) {
  Emu m;
  if (make(m, mesh, law, order)) return -1;
  m.to_soa(U);
  StageArgs a = m.base();
  a.Uin = m.Un.data();
  a.R = m.R.data();
  a.scheme = MODE_RESIDUAL;
  const int lw = m.S.law;
  int r;
  if (lw == LAW_SCALAR) r = element_residual_law<ScalarLawD>(a);
  else if (lw == LAW_EULER2D && a.law.flux == FLUX_EXACT) r = element_residual_law<EulerExactLawD>(a);
  else if (lw == LAW_EULER2D) r = element_residual_law<EulerLawD>(a);
  else if (lw == LAW_MAXWELL2D) r = element_residual_law<Maxwell2dLawD>(a);
  else if (lw == LAW_BURGERS2D) r = element_residual_law<Burgers2DLawD>(a);
  else return -3;
  if (r) return -2;
  m.from_soa(m.R.data(), R);
  return 0;
}

// ScalarPb's rotating velocity (2 pi y, -2 pi x) (Lilia.cc:66-82)
static void rotating_field(const double* x, double, double* v, void*) {
  const double pi = 3.14159265358979323846;
  v[0] = 2. * pi * x[1];
  v[1] = -2. * pi * x[0];
  v[2] = 0.0;
}

// the residual of the scalar law with the rotating velocity field: the point
// velocity table built as dg_set_velocity_field builds it
extern "C" int emu_tri_residual_rotating(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int form,
                                         const double* U, double* R) {
  const int fp32 = form & 1, shape = (form >> 1) & 1;
  Emu m;
  if (make(m, mesh, law, order)) return -1;
  std::vector<double> pv;
  if (!tri_point_field(m.S.xyz, m.S.perm, order, 2, rotating_field, nullptr, pv)) return -3;
  m.to_soa(U);
  StageArgs a = m.base();
  a.pvel = pv.data();
  a.Uin = m.Un.data();
  a.R = m.R.data();
  a.scheme = MODE_RESIDUAL;
  if (run_variant(a, m.S.law, (fp32 ? 1 : 0) | shape << 2)) return -2;
  m.from_soa(m.R.data(), R);
  return 0;
}

// nsteps steps of the given scheme in place, with the stage / buffer sequence of
// the library (dg_capi.cc enqueue_step); resid: last step's residual norm
extern "C" int emu_tri_steps(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int form, int scheme,
                             double dt, int nsteps, double* U, double* resid) {
  // form: bit 0 the FP32 form, bit 1 the 8-wave tile, bit 2 (FP32 form) the
  // library's step sequence: the first stage reads the state rounded to FP32
  // (the FP32-input kernel), the last stage writes that rounding beside U^n
  const int fp32 = form & 1, shape = (form >> 1) & 1, m32 = fp32 && ((form >> 2) & 1);
  Emu m;
  if (make(m, mesh, law, order)) return -1;
  m.to_soa(U);
  const int ns = scheme == 0 ? 1 : scheme == 1 ? 2 : scheme == 2 ? 3 : 4;
  if (fp32 && ns == 1) return -3;
  std::vector<float> M32;
  if (m32) {
    M32.resize(m.Un.size());
    for (size_t i = 0; i < M32.size(); i++) M32[i] = float(m.Un[i]);
  }
  void* bufA = fp32 ? (void*)m.A32.data() : (void*)m.A64.data();
  void* bufB = fp32 ? (void*)m.B32.data() : (void*)m.B64.data();
  for (int s = 0; s < nsteps; s++) {
    StageArgs a = m.base();
    a.dt = dt;
    a.scheme = scheme;
    for (int st = 0; st < ns; st++) {
      a.stage = st;
      const bool first = st == 0, last = st == ns - 1;
      a.Uin = first ? (m32 ? (const void*)M32.data() : (const void*)m.Un.data()) : ((st % 2) ? bufA : bufB);
      a.Uout = last ? (void*)m.Un.data() : ((st % 2) ? bufB : bufA);
      if (first && last) a.Uout = bufA;
      a.Uout32 = (last && m32) ? M32.data() : nullptr;
      a.resid_part = last ? m.part.data() : nullptr;
      const int variant = (fp32 ? (first ? (m32 ? 2 : 1) : (last ? 3 : 2)) : 0) | shape << 2;
      if (run_stage(m, a, variant)) return -2;
    }
    if (ns == 1) m.Un = m.A64;
  }
  double r = 0;
  for (int i = 0; i < (m.S.E + TILE_ELEMS - 1) / TILE_ELEMS; i++) r += m.part[i];
  if (resid) *resid = r;
  m.from_soa(m.Un.data(), U);
  return 0;
}

// nsteps local-time-stepping steps of size h in place, with the planning and
// launch sequence of the library (dg_lts_setup, enqueue_lts_step): levels by
// cell size (max_levels), elements grouped by class, the stage kernel over each
// class's range; levels (caller order, may be NULL) receives the levels
extern "C" int emu_tri_lts_steps(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int max_levels,
                                 double h, int nsteps, double* U, int32_t* levels) {
  const int E = mesh->n_elems;
  std::vector<double> size(E);
  for (int e = 0; e < E; e++) size[e] = tri_cell_size(mesh->xyz + (size_t)e * 6);
  Emu m0;
  if (make(m0, mesh, law, order)) return -1;
  std::vector<int32_t> lev, m1, m2, cid;
  std::vector<LtsClass> cls;
  const int nlev = lts_levels(E, size.data(), m0.S.faces, max_levels, lev);
  lts_reach(E, m0.S.faces, lev, m1, m2);
  lts_classify(E, lev, m1, m2, cid, cls);
  Emu m;
  if (make(m, mesh, law, order, cid.data())) return -1;
  for (auto& k : cls) k.e0 = k.e1 = -1;
  for (int i = 0; i < E; i++) {
    LtsClass& k = cls[cid[m.S.perm[i]]];
    if (k.e0 < 0) k.e0 = i;
    else if (k.e1 != i) return -4;
    k.e1 = i + 1;
  }
  if (levels) std::memcpy(levels, lev.data(), (size_t)E * 4);
  m.to_soa(U);
  LtsPlan plan;
  lts_plan(nlev, cls, plan);
  const int nrow = m.S.NF * m.S.NP;
  for (int s = 0; s < nsteps; s++) {
    for (int i = 0; i < plan.S; i++) {
      const bool odd = i & 1;
      // the stage launch: every tile as the kernel sets it up from its entry
      for (int k = plan.off[i]; k < plan.off[i + 1]; k++) {
        const LtsTile& tl = plan.tiles[k];
        StageArgs a = m.base();
        a.scheme = SCHEME_LTS;
        a.e0 = tl.e0;
        a.e1 = tl.e1;
        a.stage = tl.flags;
        a.dt = h / (double)(1 << tl.b);
        a.lts_w = tl.w;
        a.Uin = odd ? (const void*)m.A64.data() : (const void*)m.Un.data();
        a.Uout = odd ? (void*)m.Un.data() : (void*)m.A64.data();
        if (run_variant(a, m.S.law, 0)) return -2;
      }
      for (int k = plan.coff[i]; k < plan.coff[i + 1]; k++) {
        const LtsTile& tl = plan.ctiles[k];
        for (int r = 0; r < nrow; r++)
          for (int e = tl.e0; e < tl.e1; e++) m.Un[(size_t)r * E + e] += m.acc[(size_t)r * E + e];
      }
    }
  }
  m.from_soa(m.Un.data(), U);
  return 0;
}

// tangent-linear (jvp=1: dR = A dU) or adjoint (jvp=0: w = A^T z; jvp=-1, -2: the
// adjoint kernel with 1 or 2 flux tangent directions per pass) residual product
extern "C" int emu_tri_derivative(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int jvp,
                                  const double* U, const double* V, double* W) {
  Emu m;
  if (make(m, mesh, law, order)) return -1;
  const size_t n = m.Un.size();
  std::vector<double> u(n), v(n), w(n, 0.0);
  m.to_soa(U, u.data());
  m.to_soa(V, v.data());
  AdjArgs a = m.adj();
  a.U = u.data();
  a.Z = v.data();
  a.W = w.data();
  if ((jvp > 0 ? run_jvp(a, m.S.law) : run_vjp(a, m.S.law, -jvp))) return -2;
  m.from_soa(w.data(), W);
  return 0;
}

// gradient dJ/dU(0) of <seed, U(T)> over nsteps RK4 steps: the forward stage
// kernel with every stage state stored, then the fused reverse stages (the
// library's dg_adjoint_backward sequence without checkpointing)
extern "C" int emu_tri_adjoint_rk4(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, double dt,
                                   int nsteps, const double* U0, const double* seed, double* UT, double* grad) {
  Emu m;
  if (make(m, mesh, law, order)) return -1;
  const size_t n = m.Un.size();
  std::vector<double> st((size_t)nsteps * 4 * n + n);
  m.to_soa(U0, st.data());
  for (int j = 0; j < nsteps; j++) {
    double* b = st.data() + (size_t)j * 4 * n;
    StageArgs a = m.base();
    a.dt = dt;
    a.scheme = 3;
    a.Un = b;
    for (int s = 0; s < 4; s++) {
      a.stage = s;
      a.Uin = b + (size_t)s * n;
      a.Uout = s < 3 ? b + (size_t)(s + 1) * n : b + 4 * n;
      if (run_variant(a, m.S.law, 0)) return -2;
    }
  }
  m.from_soa(st.data() + (size_t)nsteps * 4 * n, UT);
  std::vector<double> lam(n), mu(n), za(n), zb(n);
  m.to_soa(seed, lam.data());
  const int BS = m.S.NF * m.S.NP;
  for (int j = nsteps - 1; j >= 0; j--) {
    for (int r = 0; r < BS; r++)
      for (int e = 0; e < m.S.E; e++) za[(size_t)r * m.S.E + e] = 1.0 / 6.0 * dt / m.S.detj[e] * lam[(size_t)r * m.S.E + e];
    AdjArgs a = m.adj();
    a.mode = ADJ_RK4;
    a.dt = dt;
    a.lam = lam.data();
    a.mu = mu.data();
    double* zin = za.data();
    double* zout = zb.data();
    for (int s = 3; s >= 0; s--) {
      a.stage = s;
      a.U = st.data() + ((size_t)j * 4 + s) * n;
      a.Z = zin;
      a.znext = zout;
      if (run_vjp(a, m.S.law)) return -2;
      std::swap(zin, zout);
    }
    std::swap(lam, mu);
  }
  m.from_soa(lam.data(), grad);
  return 0;
}

// ---------------------------------------------------------------------------
// mixed orders, limiter, sensor: the element functions of dg_mixed.h per element

namespace {

struct MixEmu {
  TriSetup S;
  std::vector<double> U, Un, acc, R, A, B;
  std::vector<uint8_t> flags, orders;
  double midv[3][15], pinv[2][3];
  void to_soa(const double* Uc, double* dst) const {
    for (int i = 0; i < S.E; i++) {
      const int p = S.order[i], n = S.NF * np_tri(p), Ep = S.bstart[p + 1] - S.bstart[p];
      for (int j = 0; j < n; j++)
        dst[S.boff[p] + (long)j * Ep + (i - S.bstart[p])] = Uc[S.coff[S.perm[i]] + j];
    }
  }
  void from_soa(const double* src, double* Uc) const {
    for (int i = 0; i < S.E; i++) {
      const int p = S.order[i], n = S.NF * np_tri(p), Ep = S.bstart[p + 1] - S.bstart[p];
      for (int j = 0; j < n; j++)
        Uc[S.coff[S.perm[i]] + j] = src[S.boff[p] + (long)j * Ep + (i - S.bstart[p])];
    }
  }
  MixedArgs args() const {
    MixedArgs m{};
    m.E = S.E;
    for (int p = 0; p < 6; p++) {
      m.bstart[p] = S.bstart[p];
      m.boff[p] = S.boff[p];
      m.moff[p] = S.moff[p];
      for (int q = 0; q < 6; q++) m.ftoff[p][q] = S.ftoff[p][q];
    }
    m.tab = S.mtab.data();
    m.glw = S.glw.data();
    m.mvtx = S.mvtx.data();
    m.el.mg = S.mg.data();
    m.el.detj = S.detj.data();
    m.el.size = S.size.data();
    m.el.link = S.link.data();
    m.law = S.lp;
    m.ntab = S.ntab.data();
    return m;
  }
  LimiterArgs lim(double* u, double M, double nu) {
    LimiterArgs l{};
    l.E = S.E;
    for (int p = 0; p < 6; p++) {
      l.bstart[p] = S.bstart[p];
      l.boff[p] = S.boff[p];
    }
    l.el.mg = S.mg.data();
    l.el.detj = S.detj.data();
    l.el.size = S.size.data();
    l.el.link = S.link.data();
    l.law = S.lp;
    l.U = u;
    l.mvtx = S.mvtx.data();
    for (int i = 0; i < 3; i++) {
      for (int k = 0; k < 15; k++) l.midv[i][k] = midv[i][k];
      for (int j = 0; j < 2; j++) l.pinv[j][i] = pinv[j][i];
    }
    l.tvb_m = M;
    l.nu = nu;
    l.flags = flags.data();
    l.tab = S.mtab.data();
    for (int p = 0; p < 6; p++) {
      l.moff[p] = S.moff[p];
      for (int q = 0; q < 6; q++) l.ftoff[p][q] = S.ftoff[p][q];
    }
    return l;
  }
};

int mix_make(MixEmu& m, const dg_mesh_desc* mesh, const dg_law_desc* law, int order) {
  dg_options opt{};
  opt.order = order;
  std::string err;
  if (build_tri_setup(mesh, law, &opt, m.S, err) != DG_OK) return -1;
  const size_t n = (size_t)m.S.boff[5];
  m.U.assign(n, 0.0);
  m.Un.assign(n, 0.0);
  m.acc.assign(n, 0.0);
  m.R.assign(n, 0.0);
  m.A.assign(n, 0.0);
  m.B.assign(n, 0.0);
  m.flags.assign(m.S.E, 0);
  m.orders.assign(m.S.E, 0);
  // limiter constants, as the library computes them
  static const double mid[3][2] = {{0.5, 0.0}, {0.5, 0.5}, {0.0, 0.5}};
  double M[3][2], phi[64], xi[3] = {0, 0, 0};
  for (int l = 0; l < 3; l++) {
    xi[0] = mid[l][0];
    xi[1] = mid[l][1];
    basis(2, MAXP_TRI, xi, phi, nullptr);
    for (int k = 0; k < 15; k++) m.midv[l][k] = phi[k];
    M[l][0] = phi[1];
    M[l][1] = phi[2];
  }
  double A[2][2] = {{0, 0}, {0, 0}};
  for (int l = 0; l < 3; l++)
    for (int r = 0; r < 2; r++)
      for (int q = 0; q < 2; q++) A[r][q] += M[l][r] * M[l][q];
  const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
  const double Ai[2][2] = {{A[1][1] / det, -A[0][1] / det}, {-A[1][0] / det, A[0][0] / det}};
  for (int r = 0; r < 2; r++)
    for (int l = 0; l < 3; l++) m.pinv[r][l] = Ai[r][0] * M[l][0] + Ai[r][1] * M[l][1];
  return 0;
}

// one tile of the mixed-order kernels, the phases in the kernels' order; PU is
// the largest order the LDS rows hold (MAXP_TRI: the persistent launch over
// every bucket; P: the launch over bucket P alone)
template <int P, class LAW, int PU, bool PF>
void mix_tile(const MixedArgs& a, MixSmem<LAW, PU>& sm, int i0, double& rsum) {
  using K = TiledMixed<P, LAW, PU, PF>;
  std::vector<typename K::Regs> g(TILE_THREADS);
  for (int t = 0; t < TILE_THREADS; t++) K::phaseA(a, sm, t, i0, g[t]);
  for (int q0 = 0; q0 < K::NQ; q0 += MIX_QC) {
    for (int t = 0; t < TILE_THREADS; t++) K::phaseB(a, sm, t, q0);
    for (int t = 0; t < TILE_THREADS; t++) K::phaseD1(a, sm, t, q0, g[t]);
  }
  for (int t = 0; t < TILE_THREADS; t++) K::phaseC1(a, sm, t, i0, g[t]);
  for (int t = 0; t < TILE_THREADS; t++) K::phaseC2(a, sm, t);
  for (int t = 0; t < TILE_THREADS; t++) {
    K::phaseD2(a, sm, t, i0, g[t]);
    rsum += g[t].rsum;
  }
}

template <int P, class LAW, int PU, bool PF>
void mix_tiles_of(const MixedArgs& a, double& rsum) {
  std::unique_ptr<MixSmem<LAW, PU>> sm(new MixSmem<LAW, PU>());
  for (int i0 = a.bstart[P]; i0 < a.bstart[P + 1]; i0 += MIX_CHUNK) {
    std::memset(sm.get(), 0, sizeof(MixSmem<LAW, PU>));
    for (int t = 0; t < TILE_THREADS; t++) TiledMixed<P, LAW, PU, PF>::stage_tables(a, *sm, t);
    mix_tile<P, LAW, PU, PF>(a, *sm, i0, rsum);
  }
}

// 0: the tiles as the persistent launch runs them (LDS rows of the top order);
// 1: as the bucket-by-bucket launches run them (each bucket's own LDS rows)
static int g_mix_mode = 0;
extern "C" void emu_mixed_mode(int mode) { g_mix_mode = mode; }

// every chunk of 64 elements of every bucket
template <class LAW>
void mix_stage_all(const MixedArgs& a, double& rsum) {
  if (g_mix_mode == 1) {
    mix_tiles_of<4, LAW, 4, true>(a, rsum);
    mix_tiles_of<3, LAW, 3, true>(a, rsum);
    mix_tiles_of<2, LAW, 2, true>(a, rsum);
    mix_tiles_of<1, LAW, 1, true>(a, rsum);
  } else {
    mix_tiles_of<4, LAW, MAXP_TRI, false>(a, rsum);
    mix_tiles_of<3, LAW, MAXP_TRI, false>(a, rsum);
    mix_tiles_of<2, LAW, MAXP_TRI, false>(a, rsum);
    mix_tiles_of<1, LAW, MAXP_TRI, false>(a, rsum);
  }
}

void mix_stage(const MixedArgs& a, int law, double& rsum) {
  if (law == LAW_SCALAR)
    mix_stage_all<ScalarLawD>(a, rsum);
  else if (law == LAW_MAXWELL2D)
    mix_stage_all<Maxwell2dLawD>(a, rsum);
  else if (law == LAW_BURGERS2D)
    mix_stage_all<Burgers2DLawD>(a, rsum);
  else if (a.law.char_bc)
    mix_stage_all<EulerCharLawD>(a, rsum);
  else
    mix_stage_all<EulerLawD>(a, rsum);
}

void mix_limit(const LimiterArgs& l, int law) {
  for (int i = 0; i < l.E; i++) {
    if (law == LAW_SCALAR)
      limit_element<ScalarLawD>(l, i);
    else if (l.law.char_bc)
      limit_element<EulerCharLawD>(l, i);
    else
      limit_element<EulerLawD>(l, i);
  }
}

}  // namespace

// residual of a mesh with per-element orders (mesh->order) through the mixed kernel
extern "C" int emu_mixed_residual(const dg_mesh_desc* mesh, const dg_law_desc* law, const double* U, double* R) {
  MixEmu m;
  if (mix_make(m, mesh, law, 1)) return -1;
  m.to_soa(U, m.U.data());
  MixedArgs a = m.args();
  a.Uin = m.U.data();
  a.R = m.R.data();
  a.scheme = MODE_RESIDUAL;
  double r = 0;
  mix_stage(a, m.S.law, r);
  m.from_soa(m.R.data(), R);
  return 0;
}

// nsteps of a scheme with the library's buffer sequence, the limiter after every
// stage when lim_on; flags: the last limiter pass (caller order)
extern "C" int emu_mixed_steps(const dg_mesh_desc* mesh, const dg_law_desc* law, int scheme, double dt, int nsteps,
                               int lim_on, double M, double nu, double* U, double* resid, uint8_t* flags) {
  MixEmu m;
  if (mix_make(m, mesh, law, 1)) return -1;
  m.to_soa(U, m.Un.data());
  const int ns = scheme == 0 ? 1 : scheme == 1 ? 2 : scheme == 2 ? 3 : 4;
  double r = 0;
  for (int st = 0; st < nsteps; st++) {
    MixedArgs a = m.args();
    a.Un = m.Un.data();
    a.acc = m.acc.data();
    a.dt = dt;
    a.scheme = scheme;
    r = 0;
    for (int s = 0; s < ns; s++) {
      a.stage = s;
      const bool first = s == 0, last = s == ns - 1;
      a.Uin = first ? m.Un.data() : ((s % 2) ? m.A.data() : m.B.data());
      a.Uout = last ? m.Un.data() : ((s % 2) ? m.B.data() : m.A.data());
      if (first && last) a.Uout = m.A.data();
      mix_stage(a, m.S.law, r);
      if (lim_on) mix_limit(m.lim(a.Uout, M, nu), m.S.law);
    }
    if (ns == 1) m.Un = m.A;
  }
  if (resid) *resid = r;
  m.from_soa(m.Un.data(), U);
  if (flags)
    for (int i = 0; i < m.S.E; i++) flags[m.S.perm[i]] = m.flags[i];
  return 0;
}

extern "C" int emu_limit(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, double M, double nu, double* U,
                         uint8_t* flags) {
  MixEmu m;
  if (mix_make(m, mesh, law, order)) return -1;
  m.to_soa(U, m.U.data());
  mix_limit(m.lim(m.U.data(), M, nu), m.S.law);
  m.from_soa(m.U.data(), U);
  for (int i = 0; i < m.S.E; i++) flags[m.S.perm[i]] = m.flags[i];
  return 0;
}

// the non-physical recovery pass (apply) or check of dg_mixed.h over every
// element; flags (caller order) may be NULL, as in dg_advance's check after
// its steps (the element function then records bits in the status only)
extern "C" int emu_recover(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int apply, double* U,
                           uint8_t* flags, unsigned* status) {
  MixEmu m;
  if (mix_make(m, mesh, law, order)) return -1;
  m.to_soa(U, m.U.data());
  std::fill(m.flags.begin(), m.flags.end(), 0);
  RecoverArgs r{};
  r.l = m.lim(m.U.data(), 0.0, 1.0);
  r.l.flags = flags ? m.flags.data() : nullptr;
  r.apply = apply;
  unsigned st = 0;
  for (int i = 0; i < m.S.E; i++) {
    const int p = mixed_order_of(r.l, i);
    unsigned b = 0;
    if (m.S.law == LAW_EULER2D) {
      switch (p) {
        case 1: b = recover_element_p<4, true, 1>(r, i); break;
        case 2: b = recover_element_p<4, true, 2>(r, i); break;
        case 3: b = recover_element_p<4, true, 3>(r, i); break;
        default: b = recover_element_p<4, true, 4>(r, i); break;
      }
    } else {
      switch (p) {
        case 1: b = recover_element_p<1, false, 1>(r, i); break;
        case 2: b = recover_element_p<1, false, 2>(r, i); break;
        case 3: b = recover_element_p<1, false, 3>(r, i); break;
        default: b = recover_element_p<1, false, 4>(r, i); break;
      }
    }
    st |= b;
  }
  m.from_soa(m.U.data(), U);
  if (flags)
    for (int i = 0; i < m.S.E; i++) flags[m.S.perm[i]] = m.flags[i];
  if (status) *status = st;
  return 0;
}

extern "C" int emu_sensor(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, const double* U,
                          const double* s_hi, const double* s_lo, int pmin, int pmax, int32_t* orders) {
  MixEmu m;
  if (mix_make(m, mesh, law, order)) return -1;
  m.to_soa(U, m.U.data());
  SensorArgs a{};
  a.E = m.S.E;
  for (int p = 0; p < 6; p++) {
    a.bstart[p] = m.S.bstart[p];
    a.boff[p] = m.S.boff[p];
    a.s_hi[p] = p >= 1 && p <= 4 ? s_hi[p - 1] : 0.0;
    a.s_lo[p] = p >= 1 && p <= 4 ? s_lo[p - 1] : 0.0;
  }
  a.pmin = pmin;
  a.pmax = pmax;
  a.U = m.U.data();
  a.orders = m.orders.data();
  for (int i = 0; i < a.E; i++) sensor_element(a, i);
  for (int i = 0; i < m.S.E; i++) orders[m.S.perm[i]] = m.orders[i];
  return 0;
}

// ---------------------------------------------------------------------------
// tetrahedra (Maxwell): the tiled tet kernel's phases

namespace {

template <int P, int NW, bool ALIAS, typename CT = double, typename IT = double, typename OT = double>
void tet_tiles_run(const TetArgs& a, double* rsum) {
  using K = TiledTet<P, NW, ALIAS, CT, IT, OT>;
  constexpr int TH = K::THREADS;
  std::unique_ptr<typename K::Smem> sm(new typename K::Smem());
  std::vector<typename K::Regs> g(TH);
  const int ntiles = (a.E + TILE_ELEMS - 1) / TILE_ELEMS;
  for (int tile = 0; tile < ntiles; tile++) {
    std::memset(sm.get(), 0, sizeof(typename K::Smem));
    for (int t = 0; t < TH; t++) K::phaseA(a, *sm, t, tile);
    for (int t = 0; t < TH; t++) K::phaseB(a, *sm, t, tile, g[t]);
    for (int l = 0; l < 4; l++) {
      for (int t = 0; t < TH; t++) K::phaseC1(a, *sm, t, tile, l);
      for (int t = 0; t < TH; t++) K::phaseC2(a, *sm, t, tile, l);
      for (int t = 0; t < TH; t++) K::phaseLift(a, *sm, t, l, g[t]);
    }
    for (int t = 0; t < TH; t++) {
      K::phaseD(a, t, tile, g[t]);
      *rsum += g[t].rsum;
    }
  }
}

template <int NW, bool ALIAS>
int tet_run_nw(const TetArgs& a, double* rsum) {
  switch (a.P) {
    case 1: tet_tiles_run<1, NW, ALIAS>(a, rsum); return 0;
    case 2: tet_tiles_run<2, NW, ALIAS>(a, rsum); return 0;
    case 3: tet_tiles_run<3, NW, ALIAS>(a, rsum); return 0;
  }
  return -1;
}

template <int NW, bool ALIAS, typename CT, typename IT, typename OT>
int tet_run_nw_form(const TetArgs& a, double* rsum) {
  switch (a.P) {
    case 1: tet_tiles_run<1, NW, ALIAS, CT, IT, OT>(a, rsum); return 0;
    case 2: tet_tiles_run<2, NW, ALIAS, CT, IT, OT>(a, rsum); return 0;
    case 3: tet_tiles_run<3, NW, ALIAS, CT, IT, OT>(a, rsum); return 0;
  }
  return -1;
}

// the FP32 form (prec 1..3 as the device launcher's) for one shape
template <typename IT, typename OT>
int tet_run32_io(const TetArgs& a, int shape, double* rsum) {
  switch (shape) {
    case 0: return tet_run_nw_form<4, true, float, IT, OT>(a, rsum);
    case 1: return tet_run_nw_form<8, true, float, IT, OT>(a, rsum);
    case 2: return tet_run_nw_form<8, false, float, IT, OT>(a, rsum);
    case 3: return tet_run_nw_form<16, false, float, IT, OT>(a, rsum);
    case 4: return tet_run_nw_form<4, false, float, IT, OT>(a, rsum);
  }
  return -1;
}
int tet_run32(const TetArgs& a, int shape, int prec, double* rsum) {
  switch (prec) {
    case 1: return tet_run32_io<double, float>(a, shape, rsum);
    case 2: return tet_run32_io<float, float>(a, shape, rsum);
    case 3: return tet_run32_io<float, double>(a, shape, rsum);
  }
  return -1;
}

// shape as the device launcher's
int tet_run(const TetArgs& a, int shape, double* rsum) {
  switch (shape) {
    case 0: return tet_run_nw<4, true>(a, rsum);
    case 1: return tet_run_nw<8, true>(a, rsum);
    case 2: return tet_run_nw<8, false>(a, rsum);
    case 3: return tet_run_nw<16, false>(a, rsum);
    case 4: return tet_run_nw<4, false>(a, rsum);
  }
  return -1;
}

struct TetEmu {
  TetSetup S;
  std::vector<double> Un, acc, R, A, B;
  std::vector<float> tabf, A32, B32;
  void to_soa(const double* U, double* dst) const {
    const int BS = S.NF * S.NP;
    for (int i = 0; i < S.E; i++)
      for (int j = 0; j < BS; j++) dst[(size_t)j * S.E + i] = U[(size_t)S.perm[i] * BS + j];
  }
  void from_soa(const double* src, double* U) const {
    const int BS = S.NF * S.NP;
    for (int i = 0; i < S.E; i++)
      for (int j = 0; j < BS; j++) U[(size_t)S.perm[i] * BS + j] = src[(size_t)j * S.E + i];
  }
  TetArgs args() const {
    TetArgs a{};
    a.E = S.E;
    a.P = S.P;
    a.tab = S.tab.data();
    a.tabf = tabf.data();
    a.mg = S.mg.data();
    a.detj = S.detj.data();
    a.link = S.link.data();
    a.law = S.lp;
    return a;
  }
};

int tet_make(TetEmu& m, const dg_mesh_desc* mesh, const dg_law_desc* law, int order) {
  dg_options opt{};
  opt.order = order;
  opt.precision = DG_PREC_FP64;
  std::string err;
  if (build_tet_setup(mesh, law, &opt, m.S, err) != DG_OK) return -1;
  const size_t n = (size_t)m.S.E * m.S.NF * m.S.NP;
  m.Un.assign(n, 0.0);
  m.acc.assign(n, 0.0);
  m.R.assign(n, 0.0);
  m.A.assign(n, 0.0);
  m.B.assign(n, 0.0);
  m.A32.assign(n, 0.0f);
  m.B32.assign(n, 0.0f);
  m.tabf.assign(m.S.tab.begin(), m.S.tab.end());
  return 0;
}

}  // namespace

extern "C" int emu_tet_residual(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int shape,
                                const double* U, double* R) {
  TetEmu m;
  if (tet_make(m, mesh, law, order)) return -1;
  m.to_soa(U, m.Un.data());
  TetArgs a = m.args();
  a.Uin = m.Un.data();
  a.R = m.R.data();
  a.scheme = MODE_RESIDUAL;
  double r = 0;
  if (tet_run(a, shape, &r)) return -2;
  m.from_soa(m.R.data(), R);
  return 0;
}

extern "C" int emu_tet_steps(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int shape, int scheme,
                             double dt, int nsteps, double* U, double* resid) {
  TetEmu m;
  if (tet_make(m, mesh, law, order)) return -1;
  m.to_soa(U, m.Un.data());
  const int ns = scheme == 0 ? 1 : scheme == 1 ? 2 : scheme == 2 ? 3 : 4;
  double r = 0;
  for (int st = 0; st < nsteps; st++) {
    TetArgs a = m.args();
    a.Un = m.Un.data();
    a.acc = m.acc.data();
    a.dt = dt;
    a.scheme = scheme;
    r = 0;
    for (int s = 0; s < ns; s++) {
      a.stage = s;
      const bool first = s == 0, last = s == ns - 1;
      a.Uin = first ? m.Un.data() : ((s % 2) ? m.A.data() : m.B.data());
      a.Uout = last ? m.Un.data() : ((s % 2) ? m.B.data() : m.A.data());
      if (first && last) a.Uout = m.A.data();
      double rs = 0;
      if (tet_run(a, shape, &rs)) return -2;
      if (last) r = rs;
    }
    if (ns == 1) m.Un = m.A;
  }
  if (resid) *resid = r;
  m.from_soa(m.Un.data(), U);
  return 0;
}

// 1: emu_tet_steps32 runs the library's step sequence, whose first stage reads
// the step-start state rounded to FP32 (written by the step before's last
// stage); 0: the first stage reads the FP64 state
static int g_tet_m32 = 0;
extern "C" void emu_tet_m32(int on) { g_tet_m32 = on; }

// RK steps of the tetrahedron kernel's FP32 form with the library's buffer
// sequence: stage 1 reads the step-start state and writes an FP32 copy, the
// middle stages run FP32 to FP32, the last writes the FP64 state
extern "C" int emu_tet_steps32(const dg_mesh_desc* mesh, const dg_law_desc* law, int order, int shape, int scheme,
                               double dt, int nsteps, double* U) {
  TetEmu m;
  if (tet_make(m, mesh, law, order)) return -1;
  m.to_soa(U, m.Un.data());
  const bool m32 = g_tet_m32 != 0;
  std::vector<float> M32(m32 ? m.Un.size() : 0);
  for (size_t i = 0; i < M32.size(); i++) M32[i] = float(m.Un[i]);
  const int ns = scheme == 1 ? 2 : scheme == 2 ? 3 : 4;
  if (scheme == 0) return -3;
  for (int st = 0; st < nsteps; st++) {
    TetArgs a = m.args();
    a.Un = m.Un.data();
    a.acc = m.acc.data();
    a.dt = dt;
    a.scheme = scheme;
    for (int s = 0; s < ns; s++) {
      a.stage = s;
      const bool first = s == 0, last = s == ns - 1;
      a.Uin = first ? (m32 ? (const void*)M32.data() : (const void*)m.Un.data())
                    : (const void*)((s % 2) ? m.A32.data() : m.B32.data());
      a.Uout = last ? (void*)m.Un.data() : (void*)((s % 2) ? m.B32.data() : m.A32.data());
      a.Uout32 = (last && m32) ? M32.data() : nullptr;
      double rs = 0;
      if (tet_run32(a, shape, first ? (m32 ? 2 : 1) : (last ? 3 : 2), &rs)) return -2;
    }
  }
  m.from_soa(m.Un.data(), U);
  return 0;
}
