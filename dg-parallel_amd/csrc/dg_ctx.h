// The context behind the opaque dg_ctx handle of the C ABI: host copies of the
// caller's mesh, the element tables' bucket layout and the device buffers.
// Private to the library's sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/dg_capi.h"
#include "dg_laws.h"
#include "dg_lts.h"
#include "dg_setup.h"

using namespace dgp;

struct dg_ctx {
  std::string err;
  int dim = 2, E = 0, Nf = 0, P = 0, NP = 0;
  int law = 0, prec = 0, use_graph = 0, device = 0;
  double hmin = 0.0;  // tetrahedra: shortest edge (the Maxwell dt needs no state)
  int kernel = 0;  // 0: tiled LDS-staged stage kernel, 1: one element per work-item
  // launch schedule (dg_set_schedule / dg_autotune): tile shape of the tiled
  // kernels, workgroups of the persistent mixed-order launch (0: default)
  int shape = 0, mix_grid = 0;
  size_t part_cap = 0;  // doubles in part
  // mixed orders: caller topology and geometry (rebucketing keeps them), the
  // capacity of the state buffers (every element at the top order) and of the
  // chunk list, and the old-index map of a rebucketing
  TriTopo topo;
  size_t state_cap = 0;
  int chunk_cap = 0;
  int32_t* remap = nullptr;
  int adj_nd = 0;  // adjoint kernel: flux tangent directions per pass (0: default)
  LawParams lp{};
  size_t ndof = 0;
  std::vector<int32_t> perm;  // internal index -> caller element
  std::vector<int32_t> faces; // (eL, lL, eR, lR) caller numbering
  std::vector<double> xyz;    // caller corner coordinates [E][3][2]
  // order buckets (a uniform mesh is one bucket at offset 0)
  bool mixed = false;
  int bstart[6] = {0, 0, 0, 0, 0, 0};
  long boff[6] = {0, 0, 0, 0, 0, 0};
  std::vector<int8_t> order;  // internal order per element
  std::vector<long> coff;     // caller element -> first coefficient (E+1)
  int moff[6] = {0, 0, 0, 0, 0, 0};
  int ftoff[6][6];
  int nchunks = 0;
  // what dg_create was given, kept to rebuild the buckets when orders change
  std::vector<int32_t> m_vids, m_bc, m_order;
  dg_law_desc m_law{};
  dg_options m_opt{};
  // limiter (applied after every RK stage when on)
  int lim_on = 0;
  double lim_m = 0.0, lim_nu = 1.0;
  double lim_midv[3][15], lim_pinv[2][3];
  hipStream_t stream = nullptr;
  // mixed orders with the limiter: three side streams and fork/join events that
  // let the order buckets' limiter launches run side by side (also when captured)
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
  // device memory
  double *Un = nullptr, *acc = nullptr, *A64 = nullptr, *B64 = nullptr, *R = nullptr;
  float *A32 = nullptr, *B32 = nullptr;
  // FP32 form (tetrahedra, uniform-order triangles): the step-start state rounded to FP32, written by
  // each step's last stage and read by the next step's first stage
  float* M32 = nullptr;
  double *mg = nullptr, *detj = nullptr, *size = nullptr, *tab = nullptr, *vtx = nullptr, *part = nullptr;
  // mixed orders: second geometry buffers, the targets of re-bucketing's device
  // gather (swapped with mg, detj, size after it)
  double *mg2 = nullptr, *detj2 = nullptr, *size2 = nullptr;
  float* tabf = nullptr;
  int32_t* link = nullptr;
  double *mtab = nullptr, *glw = nullptr, *mvtx = nullptr, *ntab = nullptr;
  int* chunks = nullptr;
  uint8_t *flags = nullptr, *sens = nullptr;
  // non-physical recovery (dg_set_recovery) after every stage when on; per-element
  // flags of the recovery passes of the last dg_advance (or of dg_recover), and
  // two status words: [0] the recovery passes, [1] the check after the steps
  int rec_on = 0;
  // the scalar law's velocity at every quadrature point (dg_set_velocity_field), or null
  double* pvel = nullptr;
  // local time stepping (dg_lts_setup): levels (caller numbering), the element
  // classes' internal ranges (the layout groups each class: lts_grp, caller
  // numbering), and 2^level per internal element on the device (the time step
  // of the coarsest level)
  int lts_nlev = 0;
  std::vector<int32_t> lts_grp, lts_level;
  std::vector<LtsClass> lts_cls;
  LtsPlan lts_plan;
  LtsTile *lts_tiles = nullptr, *lts_ctiles = nullptr;  // the plan's tile lists (device)
  double* dtscale = nullptr;
  // hanging nodes (faces of half an edge): the elements that own such faces,
  // their records (HangArgs) and the half-face basis tables of order P
  int hang_n = 0;
  int32_t *hang_elem = nullptr, *hang_rstart = nullptr, *hang_rec = nullptr;
  double* phih = nullptr;
  // h-adaptivity (dg_adapt_h): per current element its refinement level, its
  // family (index into hfam; -1: an element of the mesh given to dg_create) and
  // its place among the family's four children; each family keeps its parent's
  // corners, tags and own history; the midpoint id of every edge ever split
  // (sorted end ids -> id), and the next free vertex id
  struct HFam {
    int32_t vids[3], bc[3], level, fam, ci;
    double xyz[6];
  };
  std::vector<int32_t> h_level, h_fam, h_ci;
  std::vector<HFam> hfam;
  std::map<uint64_t, int32_t> h_mid;
  int32_t h_nextid = -1;
  uint8_t* npflags = nullptr;
  unsigned* status = nullptr;
  std::vector<void*> allocs;
  // captured step graphs: gexec holds several steps, gexec1 one step
  hipGraphExec_t gexec = nullptr, gexec1 = nullptr;
  int g_scheme = -1;
  double g_dt = 0.0;
  // discrete adjoint (FP64, RK4): checkpoints every adj_every steps, the stage
  // states of one segment, and the adjoint work vectors
  double *ckpt = nullptr, *seg = nullptr, *lam = nullptr, *mu = nullptr, *za = nullptr, *zb = nullptr,
         *wv = nullptr;
  int adj_nsteps = 0, adj_every = 0, adj_nck = 0;
  size_t adj_ck_cap = 0, adj_seg_cap = 0;  // doubles allocated for ckpt and seg
  double adj_dt = 0.0;
  bool adj_ready = false;
  std::vector<double> adj_seed;  // dJ/dU(T) of the density-misfit objective, caller layout
  float adj_fwd_ms = 0.f, adj_bwd_ms = 0.f;
};

inline dg_status fail(dg_ctx* c, dg_status s, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return s;
}

#define HIPCHECK(c, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return fail((c), DG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

inline bool ready(const dg_ctx* c) { return c && c->stream && c->Un; }

// a new triangle mesh of the same order for the context (dg_adapt_h)
dg_status ctx_remesh(dg_ctx* c, const std::vector<int32_t>& vids, const std::vector<double>& xyz,
                     const std::vector<int32_t>& bc, const std::vector<double>& U);

// new per-element orders (caller numbering) for a triangle context, carrying
// the state over (raised orders get zero modes, lowered orders drop the top ones)
dg_status ctx_rebucket(dg_ctx* c, const std::vector<int32_t>& neword);

// per-field L1, L2, Linf of (state - exact field) (outputs may be NULL)
dg_status ctx_error_norms(dg_ctx* c, int field, const double* par, double t, double* l1, double* l2, double* linf);
