// C ABI implementation: host-side table construction (FP64), device buffers,
// kernel launches and hipGraph replay of Runge-Kutta steps.
//
// Host setup replaces what DGAnalysis's constructor builds per cell and per face
// (DGAnalysis.cc:133-302, DGCell.cc:127-182, 423-555) with structure-of-arrays
// tables: per element |J| J^-1, |J|, size and three face links; per order a
// block of basis values at quadrature points shared by all elements.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/dg_capi.h"
#include "dg_adjoint.h"
#include "dg_ctx.h"
#include "dg_fields.h"
#include "dg_kernels.h"
#include "dg_mixed.h"
#include "dg_tet.h"
#include "dg_setup.h"
#include "dg_tables.h"

using namespace dgp;


template <typename T>
static dg_status dalloc(dg_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, DG_ERR_OOM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
  }
  c->allocs.push_back(q);
  *p = (T*)q;
  return DG_OK;
}

// free one buffer obtained from dalloc (synchronises the device)
static void release(dg_ctx* c, void* p) {
  for (size_t i = 0; i < c->allocs.size(); i++)
    if (c->allocs[i] == p) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipFree(p);
      c->allocs.erase(c->allocs.begin() + i);
      return;
    }
}

// the captured step graphs hold the context's buffers, launch schedule and
// stage options: dropped whenever one of them changes
static void drop_graphs(dg_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->gexec1) (void)hipGraphExecDestroy(c->gexec1);
  c->gexec = c->gexec1 = nullptr;
  c->g_scheme = -1;
}

static void free_all(dg_ctx* c) {
  drop_graphs(c);
  for (int i = 0; i < 3; i++) {
    if (c->side[i]) (void)hipStreamDestroy(c->side[i]);
    if (c->ev_join[i]) (void)hipEventDestroy(c->ev_join[i]);
    c->side[i] = nullptr;
    c->ev_join[i] = nullptr;
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  c->ev_fork = nullptr;
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

static int nfields(int law) {
  switch (law) {
    case DG_LAW_SCALAR: return 1;
    case DG_LAW_EULER2D: return 4;
    case DG_LAW_MAXWELL3D: return 6;
    case DG_LAW_MAXWELL2D: return 3;
    case DG_LAW_BURGERS2D: return 1;
  }
  return -1;
}


extern "C" {

const char* dg_last_error(const dg_ctx* c) { return c ? c->err.c_str() : "null context"; }
size_t dg_ndof(const dg_ctx* c) { return c ? c->ndof : 0; }
int dg_nfaces(const dg_ctx* c) { return c ? (int)(c->faces.size() / 4) : 0; }

dg_status dg_faces(const dg_ctx* c, int32_t* out4, size_t n) {
  if (!c || !out4 || n < c->faces.size()) return DG_ERR_ARG;
  std::memcpy(out4, c->faces.data(), c->faces.size() * sizeof(int32_t));
  return DG_OK;
}

// chunks of the persistent mixed-order launch: one per MIX_CHUNK elements of a
// bucket, as (order, first internal element) pairs, in the caller position of
// their first element (each bucket runs in caller-block order, so this
// interleaves the buckets region by region: the tiles in flight together
// gather neighbours of every order from the same part of the mesh, which stays
// in L2), ties heaviest order first
static std::vector<int> mixed_chunk_list(const int* bstart, const std::vector<int32_t>& perm) {
  std::vector<std::pair<int, int>> ch;  // (order, first)
  for (int p = MAXP_TRI; p >= 1; p--)
    for (int i = bstart[p]; i < bstart[p + 1]; i += MIX_CHUNK) ch.push_back({p, i});
  std::stable_sort(ch.begin(), ch.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
    return perm[a.second] < perm[b.second];
  });
  std::vector<int> chunks;
  chunks.reserve(ch.size() * 2);
  for (auto& c : ch) {
    chunks.push_back(c.first);
    chunks.push_back(c.second);
  }
  return chunks;
}
// the most chunks any order map of E elements gives
static int max_chunks(int E) { return (E + MIX_CHUNK - 1) / MIX_CHUNK + MAXP_TRI; }

// per-workgroup partial sums / minima: one slot per workgroup of the largest launch
static size_t part_size(const TriSetup& S) {
  const int E = S.E;
  int n = std::max(tri_stage_blocks(E), tri_tiled_blocks(E));
  n = std::max(n, max_chunks(E));  // the bucket-by-bucket mixed launches: one per chunk
  n = std::max(n, mixed_dt_blocks(E));
  return (size_t)n + 1;
}

// tetrahedra: one order, Maxwell, FP64; one order bucket
static dg_status ctx_init_tet(dg_ctx* c, const dg_mesh_desc* m, const dg_law_desc* law, const dg_options* opt) {
  TetSetup S;
  std::string err;
  dg_status st = build_tet_setup(m, law, opt, S, err);
  if (st != DG_OK) return fail(c, st, "%s", err.c_str());
  c->dim = 3;
  c->E = S.E;
  c->law = S.law;
  c->Nf = S.NF;
  c->P = S.P;
  c->NP = S.NP;
  c->prec = opt->precision;
  if (c->prec != DG_PREC_FP64 && c->prec != DG_PREC_FP32) return fail(c, DG_ERR_ARG, "unknown precision %d", c->prec);
  c->use_graph = opt->use_graph;
  c->device = opt->device;
  c->lp = S.lp;
  c->mixed = false;
  for (int p = 0; p < 6; p++) {
    c->bstart[p] = p <= S.P ? 0 : S.E;
    c->boff[p] = p <= S.P ? 0 : (long)S.E * S.NF * S.NP;
  }
  c->order.assign(S.E, (int8_t)S.P);
  c->coff.resize(S.E + 1);
  for (int e = 0; e <= S.E; e++) c->coff[e] = (long)e * S.NF * S.NP;
  c->ndof = (size_t)S.E * S.NF * S.NP;
  c->m_vids.assign(m->vids, m->vids + (size_t)S.E * 4);
  c->m_order.assign(S.E, S.P);
  c->m_law = *law;
  c->m_opt = *opt;
  c->perm = S.perm;
  c->faces = S.faces;
  c->xyz = S.xyz;
  c->hmin = 1e300;
  for (double h : S.size) c->hmin = std::min(c->hmin, h);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return fail(c, DG_ERR_HIP, "no HIP device");
  }
  if (c->device < 0 || c->device >= ndev) return fail(c, DG_ERR_ARG, "device %d out of range", c->device);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, DG_ERR_HIP, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(c, DG_ERR_HIP, "hipStreamCreate failed");
  dg_status s;
  const size_t nd = c->ndof;
  const int E = c->E;
  if ((s = dalloc(c, &c->Un, nd)) || (s = dalloc(c, &c->acc, nd)) || (s = dalloc(c, &c->R, nd)) ||
      (s = dalloc(c, &c->mg, (size_t)9 * E)) ||
      (s = dalloc(c, &c->detj, E)) || (s = dalloc(c, &c->size, E)) || (s = dalloc(c, &c->link, (size_t)4 * E)) ||
      (s = dalloc(c, &c->tab, S.tab.size())) || (s = dalloc(c, &c->tabf, S.tab.size())) ||
      (s = dalloc(c, &c->part, c->part_cap = (size_t)tet_tiles(E) + 1)) || (s = dalloc(c, &c->status, 2)))
    return s;
  // stage copies: FP64, or FP32 in the FP32 form (step-start state and accumulator stay FP64)
  if (c->prec == DG_PREC_FP64) {
    if ((s = dalloc(c, &c->A64, nd)) || (s = dalloc(c, &c->B64, nd))) return s;
  } else {
    if ((s = dalloc(c, &c->A32, nd)) || (s = dalloc(c, &c->B32, nd)) || (s = dalloc(c, &c->M32, nd))) return s;
  }
  const std::vector<float> tabf(S.tab.begin(), S.tab.end());
  if (hipMemcpy(c->mg, S.mg.data(), S.mg.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->detj, S.detj.data(), E * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->size, S.size.data(), E * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->link, S.link.data(), S.link.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->tab, S.tab.data(), S.tab.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->tabf, tabf.data(), tabf.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->Un, 0, nd * 8) != hipSuccess || hipMemset(c->acc, 0, nd * 8) != hipSuccess)
    return fail(c, DG_ERR_HIP, "table upload failed");
  return DG_OK;
}

// Faces of half an edge (meshes with hanging nodes): per owning element (internal
// numbering) its records for the hanging-face pass, and the basis at the
// Gauss points of each face's halves (the points of the order-P face rule
// mapped to [0, 1/2] and [1/2, 1] of the face, each half's points in the
// face's direction)
static dg_status hang_setup(dg_ctx* c, const TriSetup& S) {
  if (c->mixed || c->prec != DG_PREC_FP64 || c->kernel != 0)
    return fail(c, DG_ERR_UNSUPPORTED, "meshes with hanging nodes run one order in the FP64 form of the tiled kernel");
  std::vector<int32_t> elem, rstart, rec;
  std::vector<double> ph;
  build_hang_tables(S, elem, rstart, rec, ph);
  // the pass gives each element one lane per (face half, point): at most two
  // halves on each of its three faces (1-irregular meshes)
  for (size_t i = 0; i + 1 < rstart.size(); i++)
    if (rstart[i + 1] - rstart[i] > HANG_MAXREC)
      return fail(c, DG_ERR_UNSUPPORTED, "an element has more than %d faces of half an edge", HANG_MAXREC);
  dg_status s;
  if ((s = dalloc(c, &c->hang_elem, elem.size())) || (s = dalloc(c, &c->hang_rstart, rstart.size())) ||
      (s = dalloc(c, &c->hang_rec, rec.size())) || (s = dalloc(c, &c->phih, ph.size())))
    return s;
  if (hipMemcpy(c->hang_elem, elem.data(), elem.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->hang_rstart, rstart.data(), rstart.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->hang_rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->phih, ph.data(), ph.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
    return fail(c, DG_ERR_HIP, "hanging-face table upload failed");
  c->hang_n = (int)elem.size();
  return DG_OK;
}

static dg_status ctx_init(dg_ctx* c, const dg_mesh_desc* m, const dg_law_desc* law, const dg_options* opt) {
  auto bail = [&](dg_status s) { return s; };
  if (m && m->dim == 3) return ctx_init_tet(c, m, law, opt);
  TriSetup S;
  std::string err;
  dg_status st = build_tri_setup(m, law, opt, S, err, c->lts_grp.empty() ? nullptr : c->lts_grp.data());
  if (st != DG_OK) return bail(fail(c, st, "%s", err.c_str()));
  c->dim = 2;
  c->E = S.E;
  c->law = S.law;
  c->Nf = S.NF;
  c->P = S.P;
  c->NP = S.NP;
  c->prec = opt->precision;
  c->use_graph = opt->use_graph;
  c->device = opt->device;
  c->kernel = opt->reserved[0] == 1 ? 1 : 0;
  c->adj_nd = opt->reserved[1];
  if (c->adj_nd != 0 && c->adj_nd != 1 && c->adj_nd != 2 && c->adj_nd != 4)
    return bail(fail(c, DG_ERR_ARG, "options.reserved[1] (adjoint directions per pass) must be 0, 1, 2 or 4"));
  c->lp = S.lp;
  c->mixed = S.mixed;
  for (int p = 0; p < 6; p++) {
    c->bstart[p] = S.bstart[p];
    c->boff[p] = S.boff[p];
    c->moff[p] = S.moff[p];
    for (int q = 0; q < 6; q++) c->ftoff[p][q] = S.ftoff[p][q];
  }
  c->order = S.order;
  c->coff = S.coff;
  c->ndof = (size_t)S.boff[5];
  if (c->mixed && c->kernel == 1) return bail(fail(c, DG_ERR_UNSUPPORTED, "mixed orders run the persistent mixed kernel only"));
  // the one-element-per-work-item kernel is validated for the scalar and Euler
  // laws; for Maxwell2d at orders 3 and 4 it returned wrong values on the GPU
  // (its instantiation spills thousands of scalar registers into vector lanes),
  // so the other laws run the tiled kernel only
  if (c->kernel == 1 && c->lp.char_bc)
    return bail(fail(c, DG_ERR_UNSUPPORTED, "the characteristic far field runs the tiled kernel only"));
  if (c->kernel == 1 && c->law != DG_LAW_SCALAR && c->law != DG_LAW_EULER2D)
    return bail(fail(c, DG_ERR_UNSUPPORTED, "the one-element-per-work-item kernel is built for the scalar and Euler laws"));
  c->m_vids.assign(m->vids, m->vids + (size_t)S.E * 3);
  c->m_bc.clear();
  if (m->bc) c->m_bc.assign(m->bc, m->bc + (size_t)S.E * 3);
  c->m_order.resize(S.E);
  for (int i = 0; i < S.E; i++) c->m_order[S.perm[i]] = S.order[i];
  c->m_law = *law;
  c->m_opt = *opt;
  c->perm = S.perm;
  c->faces = S.faces;
  c->xyz = S.xyz;
  if (S.mixed) c->topo = std::move(S.topo);
  const int E = c->E;
  std::vector<double>& h_mg = S.mg;
  std::vector<double>& h_detj = S.detj;
  std::vector<double>& h_size = S.size;
  std::vector<int32_t>& h_link = S.link;
  std::vector<double>& h_tab = S.tab;
  std::vector<float>& h_tabf = S.tabf;
  std::vector<double>& h_vtx = S.vtx;

  // ---- device ---------------------------------------------------------------------
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return bail(fail(c, DG_ERR_HIP, "no HIP device"));
  }
  if (c->device < 0 || c->device >= ndev) return bail(fail(c, DG_ERR_ARG, "device %d out of range", c->device));
  if (hipSetDevice(c->device) != hipSuccess) return bail(fail(c, DG_ERR_HIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(c, DG_ERR_HIP, "hipStreamCreate failed"));
  if (c->mixed) {
    if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess)
      return bail(fail(c, DG_ERR_HIP, "hipEventCreate failed"));
    for (int j = 0; j < 3; j++)
      if (hipStreamCreateWithFlags(&c->side[j], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&c->ev_join[j], hipEventDisableTiming) != hipSuccess)
        return bail(fail(c, DG_ERR_HIP, "side stream creation failed"));
  }
  dg_status s;
  const size_t nd = c->ndof;
  // mixed orders: state buffers for every element at the top order, so order
  // changes rebucket in place
  c->state_cap = c->mixed ? (size_t)E * c->Nf * np_tri(MAXP_TRI) : nd;
  const size_t ns = c->state_cap;
  if ((s = dalloc(c, &c->Un, ns)) || (s = dalloc(c, &c->acc, ns)) || (s = dalloc(c, &c->R, ns)) ||
      (s = dalloc(c, &c->mg, (size_t)4 * E)) || (s = dalloc(c, &c->detj, E)) || (s = dalloc(c, &c->size, E)) ||
      (s = dalloc(c, &c->link, (size_t)3 * E)) || (s = dalloc(c, &c->tab, h_tab.size())) ||
      (s = dalloc(c, &c->tabf, h_tabf.size())) || (s = dalloc(c, &c->vtx, h_vtx.size())) ||
      (s = dalloc(c, &c->part, c->part_cap = part_size(S))))
    return bail(s);
  if (c->mixed && ((s = dalloc(c, &c->mg2, (size_t)4 * E)) || (s = dalloc(c, &c->detj2, E)) ||
                   (s = dalloc(c, &c->size2, E))))
    return bail(s);
  if (c->prec == DG_PREC_FP64) {
    if ((s = dalloc(c, &c->A64, ns)) || (s = dalloc(c, &c->B64, ns))) return bail(s);
  } else {
    if ((s = dalloc(c, &c->A32, nd)) || (s = dalloc(c, &c->B32, nd)) || (s = dalloc(c, &c->M32, nd))) return bail(s);
  }
  // every order's tables (limiter, sensor, mixed kernels) and the chunk list of
  // the persistent mixed launch (heaviest order first)
  const std::vector<int> chunks = mixed_chunk_list(S.bstart, S.perm);
  c->nchunks = (int)chunks.size() / 2;
  c->chunk_cap = max_chunks(E);
  if (c->nchunks > c->chunk_cap) return bail(fail(c, DG_ERR_ARG, "chunk count mismatch"));
  if ((s = dalloc(c, &c->mtab, S.mtab.size())) || (s = dalloc(c, &c->glw, S.glw.size())) ||
      (s = dalloc(c, &c->mvtx, S.mvtx.size())) || (s = dalloc(c, &c->chunks, (size_t)2 * c->chunk_cap + 2)) ||
      (c->mixed && (s = dalloc(c, &c->remap, (size_t)E))) ||
      (s = dalloc(c, &c->ntab, S.ntab.size())) ||
      (s = dalloc(c, &c->flags, (size_t)E)) || (s = dalloc(c, &c->sens, (size_t)E)) ||
      (s = dalloc(c, &c->npflags, (size_t)E)) || (s = dalloc(c, &c->status, 2)))
    return bail(s);
  if (hipMemcpy(c->mtab, S.mtab.data(), S.mtab.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->glw, S.glw.data(), S.glw.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->mvtx, S.mvtx.data(), S.mvtx.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->chunks, chunks.data(), chunks.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->ntab, S.ntab.data(), S.ntab.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->flags, 0, E) != hipSuccess || hipMemset(c->npflags, 0, E) != hipSuccess)
    return bail(fail(c, DG_ERR_HIP, "table upload failed"));
  // limiter constants: phi_1, phi_2 at the edge midpoints and the least-squares
  // inverse that rebuilds (U_1, U_2) from midpoint deviations
  {
    static const double mid[3][2] = {{0.5, 0.0}, {0.5, 0.5}, {0.0, 0.5}};
    double M[3][2], phi[64], xi[3] = {0, 0, 0};
    for (int l = 0; l < 3; l++) {
      xi[0] = mid[l][0];
      xi[1] = mid[l][1];
      basis(2, MAXP_TRI, xi, phi, nullptr);
      for (int k = 0; k < 15; k++) c->lim_midv[l][k] = phi[k];
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
      for (int l = 0; l < 3; l++) c->lim_pinv[r][l] = Ai[r][0] * M[l][0] + Ai[r][1] * M[l][1];
  }
  auto up = [&](void* d, const void* h, size_t bytes) {
    return hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up(c->mg, h_mg.data(), h_mg.size() * 8) || !up(c->detj, h_detj.data(), E * 8) ||
      !up(c->size, h_size.data(), E * 8) || !up(c->link, h_link.data(), h_link.size() * 4) ||
      !up(c->tab, h_tab.data(), h_tab.size() * 8) || !up(c->tabf, h_tabf.data(), h_tabf.size() * 4) ||
      !up(c->vtx, h_vtx.data(), h_vtx.size() * 8))
    return bail(fail(c, DG_ERR_HIP, "table upload failed"));
  if (hipMemset(c->Un, 0, nd * 8) != hipSuccess || hipMemset(c->acc, 0, nd * 8) != hipSuccess)
    return bail(fail(c, DG_ERR_HIP, "memset failed"));
  if (!S.topo.hang.empty() && (s = hang_setup(c, S))) return bail(s);
  return DG_OK;
}

dg_status dg_create(const dg_mesh_desc* m, const dg_law_desc* law, const dg_options* opt, dg_ctx** out) {
  if (!out) return DG_ERR_ARG;
  *out = nullptr;
  dg_ctx* c = new (std::nothrow) dg_ctx();
  if (!c) return DG_ERR_OOM;
  dg_status s = ctx_init(c, m, law, opt);
  // on failure the context is kept (device memory released) to report the error
  if (s != DG_OK) free_all(c);
  // the allocation-free entry point's scratch address on this device (looked up
  // once per device; that entry point reports a failed lookup)
  if (s == DG_OK && hipSetDevice(c->device) == hipSuccess) (void)noalloc_scratch_init(c->device);
  *out = c;
  return s;
}

dg_status dg_destroy(dg_ctx* c) {
  if (!c) return DG_ERR_ARG;
  (void)hipSetDevice(c->device);
  free_all(c);
  delete c;
  return DG_OK;
}

static bool schedule_ok(const dg_ctx* c, const dg_schedule& s);
static void apply_schedule(dg_ctx* c, const dg_schedule& s);

// caller layout [e][f][k] (Np of each element's order, elements in caller order)
// <-> bucketed structure of arrays; visit(i, caller index, soa index)
extern "C++" {
template <class F>
static void for_each_coeff(const dg_ctx* c, F visit) {
  for (int i = 0; i < c->E; i++) {
    const int p = c->order[i], n = c->Nf * (c->dim == 3 ? np_tet(p) : np_tri(p)), Ep = c->bstart[p + 1] - c->bstart[p];
    const long cb = c->coff[c->perm[i]];
    const long sb = c->boff[p] + (i - c->bstart[p]);
    for (int j = 0; j < n; j++) visit(i, cb + j, sb + (long)j * Ep);
  }
}
}

dg_status dg_set_state(dg_ctx* c, const double* U, size_t n) {
  if (!ready(c) || !U || n != c->ndof) return fail(c, DG_ERR_ARG, "dg_set_state: need %zu doubles", c ? c->ndof : 0);
  std::vector<double> h(n);
  int bad = -1;
  for_each_coeff(c, [&](int i, long ci, long si) {
    if (!std::isfinite(U[ci]) && bad < 0) bad = c->perm[i];
    h[si] = U[ci];
  });
  if (bad >= 0) return fail(c, DG_ERR_NONFINITE, "non-finite coefficient in element %d", bad);
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipMemcpyAsync(c->Un, h.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return DG_OK;
}

static dg_status fetch_soa(dg_ctx* c, const double* dsrc, double* U) {
  std::vector<double> h(c->ndof);
  HIPCHECK(c, hipMemcpyAsync(h.data(), dsrc, c->ndof * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for_each_coeff(c, [&](int, long ci, long si) { U[ci] = h[si]; });
  return DG_OK;
}

dg_status dg_get_state(dg_ctx* c, double* U, size_t n) {
  if (!ready(c) || !U || n != c->ndof) return fail(c, DG_ERR_ARG, "dg_get_state: need %zu doubles", c ? c->ndof : 0);
  HIPCHECK(c, hipSetDevice(c->device));
  return fetch_soa(c, c->Un, U);
}

static MixedArgs mixed_args(const dg_ctx* c) {
  MixedArgs m{};
  m.E = c->E;
  for (int p = 0; p < 6; p++) {
    m.bstart[p] = c->bstart[p];
    m.boff[p] = c->boff[p];
    m.moff[p] = c->moff[p];
    for (int q = 0; q < 6; q++) m.ftoff[p][q] = c->ftoff[p][q];
  }
  m.tab = c->mtab;
  m.glw = c->glw;
  m.mvtx = c->mvtx;
  m.el.mg = c->mg;
  m.el.detj = c->detj;
  m.el.size = c->size;
  m.el.link = c->link;
  m.law = c->lp;
  m.chunks = c->chunks;
  m.ntab = c->ntab;
  m.nchunks = c->nchunks;
  return m;
}

static LimiterArgs limiter_args(const dg_ctx* c, double* U) {
  LimiterArgs l{};
  l.E = c->E;
  for (int p = 0; p < 6; p++) {
    l.bstart[p] = c->bstart[p];
    l.boff[p] = c->boff[p];
  }
  l.el.mg = c->mg;
  l.el.detj = c->detj;
  l.el.size = c->size;
  l.el.link = c->link;
  l.law = c->lp;
  l.U = U;
  l.mvtx = c->mvtx;
  for (int i = 0; i < 3; i++) {
    for (int k = 0; k < MIX_MAXNP; k++) l.midv[i][k] = c->lim_midv[i][k];
    for (int j = 0; j < 2; j++) l.pinv[j][i] = c->lim_pinv[j][i];
  }
  l.tvb_m = c->lim_m;
  l.nu = c->lim_nu;
  l.flags = c->flags;
  l.tab = c->mtab;
  for (int p = 0; p < 6; p++) {
    l.moff[p] = c->moff[p];
    for (int q = 0; q < 6; q++) l.ftoff[p][q] = c->ftoff[p][q];
  }
  return l;
}

static RecoverArgs recover_args(const dg_ctx* c, double* U, int apply, uint8_t* flags, unsigned* status) {
  RecoverArgs r{};
  r.l = limiter_args(c, U);
  r.l.flags = flags;
  r.status = status;
  r.apply = apply;
  return r;
}

// one stage kernel (uniform tiled / per-element, or the persistent mixed-order
// launch), then the limiter on its output when it is on, then the non-physical
// recovery of that output (the next residual's input) when it is on
static int launch_stage(dg_ctx* c, const StageArgs& a, int variant) {
  int r;
  if (c->dim == 3) {
    TetArgs t{};
    t.E = c->E;
    t.P = c->P;
    t.tab = c->tab;
    t.tabf = c->tabf;
    t.mg = c->mg;
    t.detj = c->detj;
    t.link = c->link;
    t.law = c->lp;
    t.Uin = a.Uin;
    t.Un = a.Un;
    t.acc = a.acc;
    t.Uout = a.Uout;
    t.R = a.R;
    t.dt = a.dt;
    t.scheme = a.scheme;
    t.stage = a.stage;
    t.resid_part = a.resid_part;
    t.Uout32 = a.Uout32;
    return launch_tet_stage(t, c->shape, variant & 3, c->stream);
  }
  if (c->mixed) {
    MixedArgs m = mixed_args(c);
    m.Uin = (const double*)a.Uin;
    m.Un = a.Un;
    m.acc = a.acc;
    m.Uout = (double*)a.Uout;
    m.R = a.R;
    m.dt = a.dt;
    m.scheme = a.scheme;
    m.stage = a.stage;
    m.resid_part = a.resid_part;
    if (c->shape == 1) {
      // bucket by bucket: orders 2..4 on the side streams, order 1 on the
      // context stream, joined before anything that follows on it
      // residual partial sums: order 4's tiles first, then 3, 2, 1
      int part0[MAXP_TRI + 1] = {0};
      for (int p = MAXP_TRI - 1; p >= 1; p--) part0[p] = part0[p + 1] + mixed_bucket_tiles(c->bstart, p + 1);
      if (hipEventRecord(c->ev_fork, c->stream) != hipSuccess) return -2;
      for (int j = 0; j < 3; j++)
        if (hipStreamWaitEvent(c->side[j], c->ev_fork, 0) != hipSuccess ||
            launch_mixed_bucket(m, c->law, MAXP_TRI - j, part0[MAXP_TRI - j], c->side[j]) ||
            hipEventRecord(c->ev_join[j], c->side[j]) != hipSuccess)
          return -2;
      r = launch_mixed_bucket(m, c->law, 1, part0[1], c->stream);
      for (int j = 0; j < 3; j++)
        if (hipStreamWaitEvent(c->stream, c->ev_join[j], 0) != hipSuccess) return -2;
    } else {
      r = launch_mixed_stage(m, c->law, c->mix_grid > 0 ? c->mix_grid : mixed_grid(c->nchunks), c->stream);
    }
  } else {
    r = c->kernel == 1 ? launch_tri_stage(a, c->law, variant, c->stream)
                       : launch_tri_stage_tiled(a, c->law, variant | c->shape << 2, c->stream);
    if (r == 0 && c->hang_n) {
      HangArgs h{};
      h.E = c->E;
      h.P = c->P;
      h.n = c->hang_n;
      h.elem = c->hang_elem;
      h.rstart = c->hang_rstart;
      h.rec = c->hang_rec;
      h.Uin = (const double*)a.Uin;
      h.Uout = (double*)a.Uout;
      h.acc = a.acc;
      h.R = a.R;
      h.tab = c->tab;
      h.phih = c->phih;
      h.el = a.el;
      h.law = c->lp;
      h.dt = a.dt;
      h.scheme = a.scheme;
      h.stage = a.stage;
      r = launch_hang(h, c->law, c->stream);
    }
  }
  if (r == 0 && c->lim_on && a.scheme != MODE_RESIDUAL) {
    const LimiterArgs la = limiter_args(c, (double*)a.Uout);
    if (c->mixed && c->side[0]) {
      // orders 2..4 on the side streams, order 1 on the context stream, joined
      // before anything that follows on the context stream
      if (hipEventRecord(c->ev_fork, c->stream) != hipSuccess) return -2;
      for (int j = 0; j < 3; j++)
        if (hipStreamWaitEvent(c->side[j], c->ev_fork, 0) != hipSuccess ||
            launch_limiter_bucket(la, c->law, 2 + j, c->side[j]) ||
            hipEventRecord(c->ev_join[j], c->side[j]) != hipSuccess)
          return -2;
      if (launch_limiter_bucket(la, c->law, 1, c->stream)) return -2;
      for (int j = 0; j < 3; j++)
        if (hipStreamWaitEvent(c->stream, c->ev_join[j], 0) != hipSuccess) return -2;
    } else {
      r = launch_limiter(la, c->law, c->stream);
    }
  }
  if (r == 0 && c->rec_on && a.scheme != MODE_RESIDUAL)
    r = launch_recover(recover_args(c, (double*)a.Uout, 1, c->npflags, c->status), c->law, c->Nf, c->stream);
  return r;
}

static int stage_blocks(const dg_ctx* c) {
  if (c->dim == 3) return tet_tiles(c->E);
  if (c->mixed && c->shape == 1) {
    int n = 0;
    for (int p = 1; p <= MAXP_TRI; p++) n += mixed_bucket_tiles(c->bstart, p);
    return n;
  }
  if (c->mixed) return c->mix_grid > 0 ? c->mix_grid : mixed_grid(c->nchunks);
  return c->kernel == 1 ? tri_stage_blocks(c->E) : tri_tiled_blocks(c->E);
}

static StageArgs base_args(dg_ctx* c) {
  StageArgs a{};
  a.E = c->E;
  a.P = c->P;
  a.tab = c->tab;
  a.tabf = c->tabf;
  a.el.mg = c->mg;
  a.el.detj = c->detj;
  a.el.size = c->size;
  a.el.link = c->link;
  a.law = c->lp;
  a.Un = c->Un;
  a.acc = c->acc;
  a.pvel = c->pvel;
  return a;
}

dg_status dg_residual(dg_ctx* c, double t, double* R_out) {
  (void)t;  // no law here has a time-dependent source or boundary value
  if (!ready(c)) return fail(c, DG_ERR_ARG, "bad context");
  HIPCHECK(c, hipSetDevice(c->device));
  StageArgs a = base_args(c);
  a.Uin = c->Un;
  a.R = c->R;
  a.scheme = MODE_RESIDUAL;
  int variant = c->prec == DG_PREC_FP64 ? 0 : 1;
  if (launch_stage(c, a, variant)) return fail(c, DG_ERR_HIP, "residual launch failed: %s", hipGetErrorString(hipGetLastError()));
  if (R_out) return fetch_soa(c, c->R, R_out);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return DG_OK;
}

// per-workgroup dt minima of the triangle kernels into part; returns their
// number, or -1 when a launch fails
static int enqueue_dt(dg_ctx* c, double cfl, double* part) {
  if (c->mixed) {
    MixedArgs m = mixed_args(c);
    m.Uin = c->Un;
    if (launch_mixed_dt(m, c->law, cfl, part, c->stream)) return -1;
    return mixed_dt_blocks(c->E);
  }
  DtArgs a{};
  a.E = c->E;
  a.P = c->P;
  a.U = c->Un;
  a.vtx = c->vtx;
  a.el.mg = c->mg;
  a.el.detj = c->detj;
  a.el.size = c->size;
  a.el.link = c->link;
  a.law = c->lp;
  a.cfl = cfl;
  a.part = part;
  a.pvel = c->pvel;
  a.dtscale = c->lts_nlev > 0 ? c->dtscale : nullptr;
  if (launch_tri_dt(a, c->law, c->stream)) return -1;
  return tri_dt_blocks(c->E);
}

// ScalarLaw's velocity FieldEvaluator (ScalarLaw.cc:29-97): the caller's field
// evaluated once, in FP64 on the host, at every volume point and every face
// point of every element (each element at its own image of a periodic face's
// point), uploaded as [NQV + 3 NQF][2][E] in the internal element order
dg_status dg_set_velocity_field(dg_ctx* c, dg_field_fn fn, void* user) {
  if (!ready(c) || !fn) return fail(c, DG_ERR_ARG, "dg_set_velocity_field: bad argument");
  if (c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "a velocity field runs meshes without hanging nodes");
  if (c->dim != 2 || c->law != DG_LAW_SCALAR)
    return fail(c, DG_ERR_UNSUPPORTED, "a velocity field serves the scalar law on triangles");
  if (c->mixed || c->kernel != 0)
    return fail(c, DG_ERR_UNSUPPORTED, "a velocity field runs the tiled kernel of uniform orders only");
  std::vector<double> h;
  if (!tri_point_field(c->xyz, c->perm, c->P, 2, fn, user, h))
    return fail(c, DG_ERR_NONFINITE, "the velocity field is not finite at a quadrature point");
  HIPCHECK(c, hipSetDevice(c->device));
  if (!c->pvel) {
    dg_status s = dalloc(c, &c->pvel, h.size());
    if (s) return s;
  }
  HIPCHECK(c, hipMemcpy(c->pvel, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  drop_graphs(c);  // captured steps change
  return DG_OK;
}

dg_status dg_compute_dt(dg_ctx* c, double cfl, double* dt_out) {
  if (!ready(c) || !dt_out || !(cfl > 0)) return fail(c, DG_ERR_ARG, "dg_compute_dt: bad argument");
  if (c->dim == 3) {
    // Maxwell's largest eigenvalue is the light speed 1 everywhere (DGCell.cc:1155-1207
    // with the TET size: the shortest edge, DGCell.cc:1230-1262)
    *dt_out = cfl * c->hmin / (2. * c->P + 1.0);
    return DG_OK;
  }
  HIPCHECK(c, hipSetDevice(c->device));
  const int nb = enqueue_dt(c, cfl, c->part);
  if (nb < 0) return fail(c, DG_ERR_HIP, "dt launch failed");
  std::vector<double> h(nb);
  HIPCHECK(c, hipMemcpyAsync(h.data(), c->part, nb * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  double dt = 1e300;
  for (double v : h) dt = std::min(dt, v);
  if (!std::isfinite(dt) || !(dt > 0)) return fail(c, DG_ERR_NONFINITE, "time step is %g", dt);
  *dt_out = dt;
  return DG_OK;
}

dg_status dg_compute_dt_noalloc(dg_ctx* c, double cfl, double* dt_dev) {
  if (!ready(c) || !dt_dev || !(cfl > 0)) return fail(c, DG_ERR_ARG, "dg_compute_dt_noalloc: bad argument");
  HIPCHECK(c, hipSetDevice(c->device));
  if (c->dim == 3) {
    // the shortest edge and the light speed 1: no reduction (as dg_compute_dt)
    if (launch_fill_one(dt_dev, cfl * c->hmin / (2. * c->P + 1.0), c->stream))
      return fail(c, DG_ERR_HIP, "dt launch failed");
    return DG_OK;
  }
  const int nb = c->mixed ? mixed_dt_blocks(c->E) : tri_dt_blocks(c->E);
  if (nb > NOALLOC_SCRATCH_DOUBLES)
    return fail(c, DG_ERR_ARG,
                "dg_compute_dt_noalloc: %d elements need %d workgroup minima; its scratch holds %d (at most %d "
                "elements); use dg_compute_dt",
                c->E, nb, NOALLOC_SCRATCH_DOUBLES, NOALLOC_SCRATCH_DOUBLES * 256);
  double* part = noalloc_scratch(c->device);
  if (!part) return fail(c, DG_ERR_HIP, "dg_compute_dt_noalloc: no scratch address for device %d", c->device);
  // every one of the nb minima is rewritten before the reduction reads it, so
  // the previous call's contents need no clearing
  if (enqueue_dt(c, cfl, part) != nb || launch_min_reduce(part, nb, dt_dev, c->stream))
    return fail(c, DG_ERR_HIP, "dt launch failed");
  return DG_OK;
}

static int nstages(int scheme) {
  switch (scheme) {
    case DG_RK_FE: return 1;
    case DG_RK_TVD2: return 2;
    case DG_RK_TVD3: return 3;
    case DG_RK4: return 4;
  }
  return -1;
}

// enqueue one full step on the stream; returns 0 on success
static int enqueue_lts_step(dg_ctx* c, double h);

// The FP32 form's steps (tetrahedra, uniform-order triangles) read their first stage from
// M32 (the step-start state rounded to FP32: the same values the FP64-input
// kernel rounds on load), so that stage runs the FP32-input kernel (fewer
// registers in flight, one more wave per SIMD); the last stage writes M32 beside
// U^n.  run_steps and dg_advance fill M32 from U^n before their steps.
static bool uses_m32(const dg_ctx* c) {
  return c->M32 && c->prec != DG_PREC_FP64 && !c->mixed && !c->hang_n;
}

static int round_state_f32(dg_ctx* c) {
  return uses_m32(c) ? launch_round_f32(c->Un, c->M32, (long)c->ndof, c->stream) : 0;
}

static int enqueue_step(dg_ctx* c, int scheme, double dt, bool want_resid) {
  if (scheme == DG_RK_TVD2_LTS) return enqueue_lts_step(c, dt);
  const int ns = nstages(scheme);
  StageArgs a = base_args(c);
  a.dt = dt;
  a.scheme = scheme;
  const bool f64 = c->prec == DG_PREC_FP64;
  const bool m32 = !f64 && ns > 1 && uses_m32(c);
  void* bufA = f64 ? (void*)c->A64 : (void*)c->A32;
  void* bufB = f64 ? (void*)c->B64 : (void*)c->B32;
  for (int st = 0; st < ns; st++) {
    a.stage = st;
    const bool first = st == 0, last = st == ns - 1;
    a.Uin = first ? (m32 ? (const void*)c->M32 : (const void*)c->Un) : ((st % 2) ? bufA : bufB);
    a.Uout = last ? (void*)c->Un : ((st % 2) ? bufB : bufA);
    if (first && last) a.Uout = bufA;  // forward Euler: never write the array being read
    a.Uout32 = (last && m32) ? c->M32 : nullptr;
    a.resid_part = (last && want_resid) ? c->part : nullptr;
    int variant;
    if (f64)
      variant = 0;
    else
      variant = first ? (m32 ? 2 : 1) : (last ? 3 : 2);
    if (launch_stage(c, a, variant)) return -1;
  }
  if (ns == 1) {
    if (hipMemcpyAsync(c->Un, c->A64, c->ndof * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) return -1;
  }
  return 0;
}

static dg_status check_scheme(dg_ctx* c, int scheme) {
  if (scheme == DG_RK_TVD2_LTS) {
    if (c->lts_nlev <= 0) return fail(c, DG_ERR_ARG, "DG_RK_TVD2_LTS needs dg_lts_setup first");
    if (c->lim_on || c->rec_on)
      return fail(c, DG_ERR_UNSUPPORTED, "DG_RK_TVD2_LTS steps apply neither the limiter nor the recovery pass");
    return DG_OK;
  }
  if (nstages(scheme) < 0) return fail(c, DG_ERR_ARG, "unknown RK scheme %d", scheme);
  if (scheme == DG_RK_FE && c->prec != DG_PREC_FP64)
    return fail(c, DG_ERR_UNSUPPORTED, "forward Euler is built for the FP64 form only");
  return DG_OK;
}

// Steps captured in one graph: one graph launch per GRAPH_STEPS steps, so the
// launch gap between graphs (a few microseconds) comes once per GRAPH_STEPS
// steps; the remainder runs the one-step graph.  Calls of fewer than
// GRAPH_MULTI_MIN steps replay the one-step graph only: capturing the longer
// graph costs more host time than it saves over a few launches (a mixed-order
// context captures again after every re-bucketing pass).
constexpr int GRAPH_STEPS = 8;
constexpr int GRAPH_MULTI_MIN = 64;

// (scheme, dt)'s steps captured into one graph, instantiated and uploaded to
// the device; advances nothing
static dg_status capture_steps(dg_ctx* c, int scheme, double dt, int n, hipGraphExec_t* out) {
  hipGraph_t g;
  HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int i = 0; i < n && !rc; i++) rc = enqueue_step(c, scheme, dt, false);
  hipError_t ec = hipStreamEndCapture(c->stream, &g);
  if (rc || ec != hipSuccess) return fail(c, DG_ERR_HIP, "graph capture failed");
  hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    *out = nullptr;
    return fail(c, DG_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ei));
  }
  HIPCHECK(c, hipGraphUpload(*out, c->stream));
  return DG_OK;
}

// the step graphs of (scheme, dt), built when missing: the one-step graph, and
// for a call of nsteps >= GRAPH_MULTI_MIN steps the GRAPH_STEPS-step graph
static dg_status ensure_graph(dg_ctx* c, int scheme, double dt, int nsteps) {
  if (!(c->gexec1 && c->g_scheme == scheme && c->g_dt == dt)) {
    drop_graphs(c);
    dg_status s = capture_steps(c, scheme, dt, 1, &c->gexec1);
    if (s) return s;
    c->g_scheme = scheme;
    c->g_dt = dt;
  }
  if (!c->gexec && nsteps >= GRAPH_MULTI_MIN) return capture_steps(c, scheme, dt, GRAPH_STEPS, &c->gexec);
  return DG_OK;
}

static dg_status run_steps(dg_ctx* c, int scheme, double dt, int nsteps) {
  if (nsteps <= 0) return DG_OK;
  if (round_state_f32(c)) return fail(c, DG_ERR_HIP, "state rounding launch failed");
  if (c->use_graph && scheme != DG_RK_FE) {
    dg_status s = ensure_graph(c, scheme, dt, nsteps);
    if (s) return s;
    int i = 0;
    if (c->gexec)
      for (; i + GRAPH_STEPS <= nsteps; i += GRAPH_STEPS) HIPCHECK(c, hipGraphLaunch(c->gexec, c->stream));
    for (; i < nsteps; i++) HIPCHECK(c, hipGraphLaunch(c->gexec1, c->stream));
  } else {
    for (int s = 0; s < nsteps; s++)
      if (enqueue_step(c, scheme, dt, false)) return fail(c, DG_ERR_HIP, "stage launch failed: %s", hipGetErrorString(hipGetLastError()));
  }
  return DG_OK;
}

// The state after dg_advance's steps: every coefficient finite and, for the
// Euler law on triangles, no element non-physical at a point its next residual
// evaluates (checked, not changed); a recovery pass that met a non-physical
// mean during the steps is reported too.
static dg_status check_after_steps(dg_ctx* c) {
  int r;
  if (c->dim == 2)
    r = launch_recover(recover_args(c, c->Un, 0, nullptr, c->status + 1), c->law, c->Nf, c->stream);
  else
    r = launch_nonfinite_scan(c->Un, (long)c->ndof, c->status + 1, c->stream);
  if (r) return fail(c, DG_ERR_HIP, "state check launch failed");
  unsigned st[2] = {0, 0};
  HIPCHECK(c, hipMemcpyAsync(st, c->status, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (st[1] & NP_NONFINITE) return fail(c, DG_ERR_NONFINITE, "the state holds NaN or Inf after the steps");
  if ((st[0] | st[1]) & NP_MEAN)
    return fail(c, DG_ERR_NONPHYSICAL, "an element's mean state is non-physical (negative density or pressure)");
  if (st[1] & NP_POINT)
    return fail(c, DG_ERR_NONPHYSICAL,
                "the state is non-physical (negative density or pressure) at quadrature points of some elements; "
                "dg_recover names them, dg_set_recovery reduces them during the steps");
  return DG_OK;
}

dg_status dg_advance(dg_ctx* c, int scheme, double t, double dt, int nsteps, double* resid_out) {
  (void)t;
  if (!ready(c) || nsteps < 0 || !(dt > 0)) return fail(c, DG_ERR_ARG, "dg_advance: bad argument");
  dg_status s = check_scheme(c, scheme);
  if (s) return s;
  HIPCHECK(c, hipSetDevice(c->device));
  // status words and the recovery flags of this call (outside any captured graph)
  HIPCHECK(c, hipMemsetAsync(c->status, 0, 2 * sizeof(unsigned), c->stream));
  if (c->dim == 2) HIPCHECK(c, hipMemsetAsync(c->npflags, 0, c->E, c->stream));
  // the first residual's input
  if (c->rec_on && launch_recover(recover_args(c, c->Un, 1, c->npflags, c->status), c->law, c->Nf, c->stream))
    return fail(c, DG_ERR_HIP, "recovery launch failed");
  if (scheme == DG_RK_TVD2_LTS && resid_out) {
    // RungeKuttaTVD2adaptive::advanceInTime returns 0 (TimeIntegrator.cc:831)
    *resid_out = 0.0;
    resid_out = nullptr;
  }
  int body = resid_out ? nsteps - 1 : nsteps;
  if ((s = run_steps(c, scheme, dt, body))) return s;
  if (resid_out && nsteps > 0) {
    if (round_state_f32(c) || enqueue_step(c, scheme, dt, true)) return fail(c, DG_ERR_HIP, "stage launch failed");
    int nb = stage_blocks(c);
    std::vector<double> h(nb);
    HIPCHECK(c, hipMemcpyAsync(h.data(), c->part, nb * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    double r = 0.0;
    for (double v : h) r += v;
    *resid_out = r;
  }
  return check_after_steps(c);
}

dg_status dg_time_advance(dg_ctx* c, int scheme, double dt, int nsteps, float* ms_out) {
  if (!ready(c) || !ms_out || nsteps <= 0) return fail(c, DG_ERR_ARG, "dg_time_advance: bad argument");
  dg_status s = check_scheme(c, scheme);
  if (s) return s;
  HIPCHECK(c, hipSetDevice(c->device));
  hipEvent_t e0, e1;
  HIPCHECK(c, hipEventCreate(&e0));
  HIPCHECK(c, hipEventCreate(&e1));
  // build the graph (if any) outside the timed region, without taking a step
  if (c->use_graph && scheme != DG_RK_FE && (s = ensure_graph(c, scheme, dt, nsteps))) return s;
  HIPCHECK(c, hipEventRecord(e0, c->stream));
  s = run_steps(c, scheme, dt, nsteps);
  HIPCHECK(c, hipEventRecord(e1, c->stream));
  HIPCHECK(c, hipEventSynchronize(e1));
  float ms = 0;
  HIPCHECK(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_out = ms;
  return s;
}

dg_status dg_set_state_device(dg_ctx* c, const double* dU) {
  if (!ready(c) || !dU) return fail(c, DG_ERR_ARG, "bad argument");
  std::vector<double> h(c->ndof);
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipMemcpy(h.data(), dU, c->ndof * 8, hipMemcpyDeviceToHost));
  return dg_set_state(c, h.data(), c->ndof);
}

dg_status dg_get_state_device(dg_ctx* c, double* dU) {
  if (!ready(c) || !dU) return fail(c, DG_ERR_ARG, "bad argument");
  std::vector<double> h(c->ndof);
  dg_status s = dg_get_state(c, h.data(), c->ndof);
  if (s) return s;
  HIPCHECK(c, hipMemcpy(dU, h.data(), c->ndof * 8, hipMemcpyHostToDevice));
  return DG_OK;
}

// tetrahedra: x(xi) = X0 + sum_c xi_c (X_{c+1} - X0)
static void tet_map(const double* X, const double* xi, double* x) {
  for (int d = 0; d < 3; d++) x[d] = X[d] + xi[0] * (X[3 + d] - X[d]) + xi[1] * (X[6 + d] - X[d]) + xi[2] * (X[9 + d] - X[d]);
}
static double tet_detj(const double* X) {
  double J[3][3];
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 3; k++) J[d][k] = X[(k + 1) * 3 + d] - X[d];
  return std::fabs(J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]));
}

// L2 projection of a field onto every element's basis with the rule of degree
// 2p (DGCell::L2ProjInitial, DGCell.cc:1458-1483; the orthonormal basis makes
// the mass matrix |J| I, so |J| cancels).  ev(x, t, val) fills the field at a
// physical point and returns 0, or an error status.
extern "C++" {
template <class Eval>
static dg_status project_with(dg_ctx* c, double t, Eval ev) {
  const int E = c->E, NF = c->Nf, dim = c->dim;
  const int pmax = dim == 2 ? MAXP_TRI : c->P;
  std::vector<QPoint> rules[MAXP_TRI + 1];
  std::vector<double> phis[MAXP_TRI + 1];
  for (int p = dim == 2 ? 1 : c->P; p <= pmax; p++) {
    rules[p] = dim == 2 ? tri_rule(2 * p) : tet_rule(2 * p);
    const int np = dim == 2 ? np_tri(p) : np_tet(p);
    phis[p].resize(rules[p].size() * np);
    for (size_t q = 0; q < rules[p].size(); q++) basis(dim, p, rules[p][q].x, &phis[p][q * np], nullptr);
  }
  std::vector<double> U(c->ndof, 0.0);
  for (int e = 0; e < E; e++) {
    const int P = dim == 2 ? c->m_order[e] : c->P, NP = dim == 2 ? np_tri(P) : np_tet(P);
    const auto& rule = rules[P];
    const auto& phi = phis[P];
    double* Ue = &U[c->coff[e]];
    // triangles weigh by |J| and divide it out after the sum, as L2ProjInitial does
    const double* X = &c->xyz[(size_t)e * (dim == 2 ? 6 : 12)];
    const double ad = dim == 2 ? std::fabs((X[2] - X[0]) * (X[5] - X[1]) - (X[4] - X[0]) * (X[3] - X[1])) : 1.0;
    for (size_t q = 0; q < rule.size(); q++) {
      double x[3] = {0.0, 0.0, 0.0}, val[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (dim == 2) {
        const double u = rule[q].x[0], v = rule[q].x[1];
        x[0] = X[0] + (X[2] - X[0]) * u + (X[4] - X[0]) * v;
        x[1] = X[1] + (X[3] - X[1]) * u + (X[5] - X[1]) * v;
      } else {
        tet_map(X, rule[q].x, x);
      }
      if (dg_status s = ev(x, t, val)) return s;
      const double jw = ad * rule[q].w;
      for (int f = 0; f < NF; f++)
        for (int k = 0; k < NP; k++) Ue[f * NP + k] += val[f] * phi[q * NP + k] * jw;
    }
    if (dim == 2) {
      const double inv = 1. / ad;
      for (int i = 0; i < NF * NP; i++) Ue[i] *= inv;
    }
  }
  return dg_set_state(c, U.data(), U.size());
}
}  // extern "C++"

dg_status dg_project(dg_ctx* c, int field, const double* par, double t) {
  if (!ready(c)) return fail(c, DG_ERR_ARG, "bad context");
  double zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (!par) par = zero;
  return project_with(c, t, [&](const double* x, double tt, double* val) -> dg_status {
    if (exact_field(field, par, c->lp, x, tt, val)) return fail(c, DG_ERR_ARG, "unknown field %d", field);
    return DG_OK;
  });
}

dg_status dg_project_fn(dg_ctx* c, dg_field_fn fn, void* user, double t) {
  if (!ready(c)) return fail(c, DG_ERR_ARG, "bad context");
  if (!fn) return fail(c, DG_ERR_ARG, "no field function");
  return project_with(c, t, [&](const double* x, double tt, double* val) -> dg_status {
    fn(x, tt, val, user);
    for (int f = 0; f < c->Nf; f++)
      if (!std::isfinite(val[f])) return fail(c, DG_ERR_ARG, "field function returned a non-finite value");
    return DG_OK;
  });
}

// Per-field L1, L2 and Linf norms of (state - exact field at t) at the points
// of the rule of degree 2(p+1) (DGCell::L1exact, densityErrorL2, LinfError:
// DGError.cc:549-580, 805-831, which take field 0); any output may be NULL
extern "C++" dg_status ctx_error_norms(dg_ctx* c, int field, const double* par, double t, double* l1, double* l2,
                                       double* linf) {
  double zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (!par) par = zero;
  const int E = c->E, NF = c->Nf, dim = c->dim, nv = dim + 1;
  std::vector<double> U(c->ndof);
  dg_status s = dg_get_state(c, U.data(), U.size());
  if (s) return s;
  const int pmax = dim == 2 ? MAXP_TRI : 3;
  std::vector<QPoint> rules[MAXP_TRI + 1];
  std::vector<double> phis[MAXP_TRI + 1];
  for (int p = 1; p <= pmax; p++) {
    rules[p] = dim == 2 ? tri_rule(std::min(2 * (p + 1), 10)) : tet_rule(std::min(2 * (p + 1), 6));
    const int np = dim == 2 ? np_tri(p) : np_tet(p);
    phis[p].resize(rules[p].size() * np);
    for (size_t q = 0; q < rules[p].size(); q++) basis(dim, p, rules[p][q].x, &phis[p][q * np], nullptr);
  }
  std::vector<double> a1(NF, 0.0), a2(NF, 0.0), ai(NF, 0.0);
  for (int e = 0; e < E; e++) {
    const int P = c->m_order[e], NP = dim == 2 ? np_tri(P) : np_tet(P);
    const auto& rule = rules[P];
    const auto& phi = phis[P];
    const double* X = &c->xyz[(size_t)e * nv * dim];
    const double ad = dim == 2 ? std::fabs((X[2] - X[0]) * (X[5] - X[1]) - (X[4] - X[0]) * (X[3] - X[1])) : tet_detj(X);
    const double* Ue = &U[c->coff[e]];
    for (size_t q = 0; q < rule.size(); q++) {
      double x[3] = {0.0, 0.0, 0.0}, val[8];
      if (dim == 2) {
        const double u = rule[q].x[0], v = rule[q].x[1];
        x[0] = X[0] + (X[2] - X[0]) * u + (X[4] - X[0]) * v;
        x[1] = X[1] + (X[3] - X[1]) * u + (X[5] - X[1]) * v;
      } else {
        tet_map(X, rule[q].x, x);
      }
      if (exact_field(field, par, c->lp, x, t, val)) return fail(c, DG_ERR_ARG, "unknown field %d", field);
      for (int f = 0; f < NF; f++) {
        double s2 = 0;
        for (int k = 0; k < NP; k++) s2 += Ue[f * NP + k] * phi[q * NP + k];
        const double d = s2 - val[f];
        a1[f] += std::fabs(d) * ad * rule[q].w;
        a2[f] += d * d * ad * rule[q].w;
        ai[f] = std::max(ai[f], std::fabs(d));
      }
    }
  }
  for (int f = 0; f < NF; f++) {
    if (l1) l1[f] = a1[f];
    if (l2) l2[f] = std::sqrt(a2[f]);
    if (linf) linf[f] = ai[f];
  }
  return DG_OK;
}

dg_status dg_l2_error(dg_ctx* c, int field, const double* par, double t, double* err) {
  if (!ready(c) || !err) return fail(c, DG_ERR_ARG, "bad argument");
  return ctx_error_norms(c, field, par, t, nullptr, err, nullptr);
}

// ---------------------------------------------------------------------------
// limiter and order selection (new designs for triangles)

dg_status dg_set_limiter(dg_ctx* c, int enabled, double tvb_m, double nu) {
  if (!ready(c) || !(tvb_m >= 0) || !(nu > 0)) return fail(c, DG_ERR_ARG, "dg_set_limiter: bad argument");
  if (enabled && c->law != DG_LAW_SCALAR && c->law != DG_LAW_EULER2D)
    return fail(c, DG_ERR_UNSUPPORTED, "the limiter is built for the scalar and Euler laws");
  if (c->dim != 2) return fail(c, DG_ERR_UNSUPPORTED, "built for triangles");
  if (enabled && c->prec != DG_PREC_FP64) return fail(c, DG_ERR_UNSUPPORTED, "the limiter runs in the FP64 form only");
  if (enabled && c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "the limiter runs meshes without hanging nodes");
  c->lim_on = enabled ? 1 : 0;
  c->lim_m = tvb_m;
  c->lim_nu = nu;
  drop_graphs(c);  // captured steps change
  return DG_OK;
}

static dg_status fetch_flags(dg_ctx* c, const uint8_t* dsrc, uint8_t* out) {
  std::vector<uint8_t> h(c->E);
  HIPCHECK(c, hipMemcpyAsync(h.data(), dsrc, c->E, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->E; i++) out[c->perm[i]] = h[i];
  return DG_OK;
}

dg_status dg_set_recovery(dg_ctx* c, int enabled) {
  if (!ready(c)) return fail(c, DG_ERR_ARG, "bad context");
  if (enabled && (c->dim != 2 || c->law != DG_LAW_EULER2D))
    return fail(c, DG_ERR_UNSUPPORTED, "recovery serves the Euler law on triangles (other laws are always physical)");
  if (enabled && c->prec != DG_PREC_FP64) return fail(c, DG_ERR_UNSUPPORTED, "recovery runs in the FP64 form only");
  if (enabled && c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "recovery runs meshes without hanging nodes");
  c->rec_on = enabled ? 1 : 0;
  drop_graphs(c);  // captured steps change
  return DG_OK;
}

dg_status dg_recover(dg_ctx* c, int apply, uint8_t* flags_out) {
  if (!ready(c)) return fail(c, DG_ERR_ARG, "bad context");
  if (c->dim != 2) return fail(c, DG_ERR_UNSUPPORTED, "built for triangles");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipMemsetAsync(c->npflags, 0, c->E, c->stream));
  HIPCHECK(c, hipMemsetAsync(c->status, 0, sizeof(unsigned), c->stream));
  if (launch_recover(recover_args(c, c->Un, apply ? 1 : 0, c->npflags, c->status), c->law, c->Nf, c->stream))
    return fail(c, DG_ERR_HIP, "recovery launch failed");
  if (flags_out) return fetch_flags(c, c->npflags, flags_out);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return DG_OK;
}

dg_status dg_recovery_flags(dg_ctx* c, uint8_t* flags_out) {
  if (!ready(c) || !flags_out) return fail(c, DG_ERR_ARG, "bad argument");
  if (c->dim != 2) return fail(c, DG_ERR_UNSUPPORTED, "built for triangles");
  HIPCHECK(c, hipSetDevice(c->device));
  return fetch_flags(c, c->npflags, flags_out);
}

dg_status dg_limit(dg_ctx* c, uint8_t* flags_out) {
  if (!ready(c)) return fail(c, DG_ERR_ARG, "bad context");
  if (c->dim != 2) return fail(c, DG_ERR_UNSUPPORTED, "built for triangles");
  if (c->prec != DG_PREC_FP64) return fail(c, DG_ERR_UNSUPPORTED, "the limiter runs in the FP64 form only");
  if (c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "the limiter runs meshes without hanging nodes");
  HIPCHECK(c, hipSetDevice(c->device));
  if (launch_limiter(limiter_args(c, c->Un), c->law, c->stream)) return fail(c, DG_ERR_HIP, "limiter launch failed");
  if (flags_out) return fetch_flags(c, c->flags, flags_out);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return DG_OK;
}

dg_status dg_limiter_flags(dg_ctx* c, uint8_t* flags_out) {
  if (!ready(c) || !flags_out) return fail(c, DG_ERR_ARG, "bad argument");
  HIPCHECK(c, hipSetDevice(c->device));
  return fetch_flags(c, c->flags, flags_out);
}

dg_status dg_get_orders(dg_ctx* c, int32_t* orders_out) {
  if (!c || !orders_out || c->m_order.empty()) return fail(c, DG_ERR_ARG, "bad argument");
  std::memcpy(orders_out, c->m_order.data(), c->m_order.size() * 4);
  return DG_OK;
}

// rebuild the buckets for new per-element orders, carrying the state over
// (orthonormal hierarchical modes: raising an order appends zero modes,
// lowering it drops the top ones)
// A mixed-order context whose new orders are still mixed: a new internal
// numbering over the kept topology, its element tables uploaded into the same
// buffers, and the state carried over on the device (the state buffers hold
// every element at the top order).  Returns false (nothing changed) when the new
// layout does not fit.
static bool rebucket_in_place(dg_ctx* c, const std::vector<int32_t>& neword, dg_status& st) {
  st = DG_OK;
  if (!c->mixed || c->topo.E != c->E || !c->remap || !c->mg2) return false;
  bool mixed = false;
  for (int e = 1; e < c->E; e++)
    if (neword[e] != neword[0]) mixed = true;
  if (!mixed) return false;  // one order again: the uniform kernels, built anew
  const int E = c->E;
  TriSetup L;
  L.E = E;
  L.NF = c->Nf;
  const std::vector<int> ord(neword.begin(), neword.end());
  tri_layout(c->topo, ord.data(), L, nullptr, false);  // the geometry moves on the device
  const std::vector<int> chunks = mixed_chunk_list(L.bstart, L.perm);
  if ((size_t)L.boff[5] > c->state_cap || (int)chunks.size() / 2 > c->chunk_cap) return false;
  // new internal element i was old internal element iold[i]
  std::vector<int32_t> at_old(E), iold(E);
  for (int i = 0; i < E; i++) at_old[c->perm[i]] = i;
  for (int i = 0; i < E; i++) iold[i] = at_old[L.perm[i]];
  RemapArgs r{};
  r.E = E;
  r.NF = c->Nf;
  for (int p = 0; p < 6; p++) {
    r.bs0[p] = c->bstart[p];
    r.bo0[p] = c->boff[p];
    r.bs1[p] = L.bstart[p];
    r.bo1[p] = L.boff[p];
  }
  r.iold = c->remap;
  r.src = c->Un;
  r.dst = c->A64;
  r.mg0 = c->mg;
  r.detj0 = c->detj;
  r.size0 = c->size;
  r.mg1 = c->mg2;
  r.detj1 = c->detj2;
  r.size1 = c->size2;
  auto up = [&](void* d, const void* h, size_t bytes) {
    return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess;
  };
  if (!up(c->remap, iold.data(), (size_t)E * 4) || launch_remap(r, c->stream) ||
      !up(c->link, L.link.data(), L.link.size() * 4) || !up(c->chunks, chunks.data(), chunks.size() * 4) ||
      hipMemsetAsync(c->flags, 0, (size_t)E, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    st = fail(c, DG_ERR_HIP, "rebucketing failed: %s", hipGetErrorString(hipGetLastError()));
    return true;
  }
  std::swap(c->Un, c->A64);
  std::swap(c->mg, c->mg2);
  std::swap(c->detj, c->detj2);
  std::swap(c->size, c->size2);
  c->perm = std::move(L.perm);
  c->order = std::move(L.order);
  c->coff = std::move(L.coff);
  for (int p = 0; p < 6; p++) {
    c->bstart[p] = L.bstart[p];
    c->boff[p] = L.boff[p];
  }
  c->ndof = (size_t)L.boff[5];
  c->nchunks = (int)chunks.size() / 2;
  c->m_order = neword;
  // the captured step holds the old buffers and chunk count
  const dg_schedule sc{c->shape, c->mix_grid, {0, 0, 0, 0, 0, 0}};
  drop_graphs(c);
  c->shape = 0;
  c->mix_grid = 0;
  if (schedule_ok(c, sc)) apply_schedule(c, sc);
  return true;
}

// a context built anew for per-element orders neword and layout groups grp
// (both caller numbering; grp may be empty), the state carried over (raised
// orders get zero modes, lowered orders drop the top ones) with the limiter,
// recovery and schedule settings
static dg_status ctx_relayout(dg_ctx* c, const std::vector<int32_t>& neword, const std::vector<int32_t>& grp) {
  std::vector<double> U(c->ndof);
  dg_status s = dg_get_state(c, U.data(), U.size());
  if (s) return s;
  std::vector<long> noff(c->E + 1, 0);
  for (int e = 0; e < c->E; e++) noff[e + 1] = noff[e] + (long)c->Nf * np_tri(neword[e]);
  std::vector<double> V(noff[c->E], 0.0);
  for (int e = 0; e < c->E; e++) {
    const int np0 = np_tri(c->m_order[e]), np1 = np_tri(neword[e]), nk = std::min(np0, np1);
    for (int f = 0; f < c->Nf; f++)
      for (int k = 0; k < nk; k++) V[noff[e] + f * np1 + k] = U[c->coff[e] + f * np0 + k];
  }
  dg_ctx* t = new (std::nothrow) dg_ctx();
  if (!t) return fail(c, DG_ERR_OOM, "out of memory");
  t->lts_grp = grp;
  dg_mesh_desc m{};
  m.dim = 2;
  m.n_elems = c->E;
  m.vids = c->m_vids.data();
  m.xyz = c->xyz.data();
  m.bc = c->m_bc.empty() ? nullptr : c->m_bc.data();
  m.order = neword.data();
  s = ctx_init(t, &m, &c->m_law, &c->m_opt);
  if (s != DG_OK) {
    c->err = t->err;
    free_all(t);
    delete t;
    return s;
  }
  t->lim_on = c->lim_on;
  t->lim_m = c->lim_m;
  t->lim_nu = c->lim_nu;
  t->rec_on = c->rec_on;
  t->adj_nd = c->adj_nd;
  const dg_schedule sc{c->shape, c->mix_grid, {0, 0, 0, 0, 0, 0}};
  if (schedule_ok(t, sc)) apply_schedule(t, sc);  // else the new buckets start from the default
  std::swap(*c, *t);
  free_all(t);
  delete t;
  return dg_set_state(c, V.data(), V.size());
}

// a context built anew on another triangle mesh of the same order (caller
// arrays: vids, xyz, bc may be empty) holding the state U (caller layout), with
// the limiter, recovery and schedule settings carried over when the new mesh
// allows them; the refinement history is the caller's to set
extern "C++" dg_status ctx_remesh(dg_ctx* c, const std::vector<int32_t>& vids, const std::vector<double>& xyz,
                                  const std::vector<int32_t>& bc, const std::vector<double>& U) {
  dg_ctx* t = new (std::nothrow) dg_ctx();
  if (!t) return fail(c, DG_ERR_OOM, "out of memory");
  dg_mesh_desc m{};
  m.dim = 2;
  m.n_elems = (int)(vids.size() / 3);
  m.vids = vids.data();
  m.xyz = xyz.data();
  m.bc = bc.empty() ? nullptr : bc.data();
  dg_options opt = c->m_opt;
  opt.order = c->P;
  dg_status s = ctx_init(t, &m, &c->m_law, &opt);
  if (s != DG_OK) {
    c->err = t->err;
    free_all(t);
    delete t;
    return s;
  }
  t->m_opt = c->m_opt;
  t->adj_nd = c->adj_nd;
  const dg_schedule sc{c->shape, c->mix_grid, {0, 0, 0, 0, 0, 0}};
  if (schedule_ok(t, sc)) apply_schedule(t, sc);
  if (!t->hang_n) {
    t->lim_on = c->lim_on;
    t->lim_m = c->lim_m;
    t->lim_nu = c->lim_nu;
    t->rec_on = c->rec_on;
  }
  // the history arrays move with the context
  t->h_level.swap(c->h_level);
  t->h_fam.swap(c->h_fam);
  t->h_ci.swap(c->h_ci);
  t->hfam.swap(c->hfam);
  t->h_mid.swap(c->h_mid);
  t->h_nextid = c->h_nextid;
  std::swap(*c, *t);
  free_all(t);
  delete t;
  return dg_set_state(c, U.data(), U.size());
}

extern "C++" dg_status ctx_rebucket(dg_ctx* c, const std::vector<int32_t>& neword) {
  dg_status st;
  if (rebucket_in_place(c, neword, st)) return st;
  // new orders end local time stepping (uniform orders only)
  return ctx_relayout(c, neword, {});
}

// ---------------------------------------------------------------------------
// local time stepping (DGAnalysis::sortCellsBySize + RungeKuttaTVD2adaptive;
// the method in dg_lts.h)

dg_status dg_lts_setup(dg_ctx* c, int max_levels, int32_t* levels_out) {
  if (!ready(c) || max_levels < 1 || max_levels > LTS_MAX_LEVELS)
    return fail(c, DG_ERR_ARG, "dg_lts_setup: bad argument (max_levels 1..%d)", LTS_MAX_LEVELS);
  if (c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "local time stepping runs meshes without hanging nodes");
  if (c->dim != 2 || c->mixed) return fail(c, DG_ERR_UNSUPPORTED, "local time stepping runs triangles of one order");
  if (c->prec != DG_PREC_FP64 || c->kernel != 0)
    return fail(c, DG_ERR_UNSUPPORTED, "local time stepping runs the FP64 form of the tiled kernel");
  if (c->pvel)
    return fail(c, DG_ERR_UNSUPPORTED, "set the velocity field after dg_lts_setup (the elements are re-laid out)");
  HIPCHECK(c, hipSetDevice(c->device));
  const int E = c->E;
  std::vector<double> size(E);
  for (int e = 0; e < E; e++) size[e] = tri_cell_size(c->xyz.data() + (size_t)e * 6);
  std::vector<int32_t> lev, m1, m2, cid;
  std::vector<LtsClass> cls;
  const int nlev = lts_levels(E, size.data(), c->faces, max_levels, lev);
  lts_reach(E, c->faces, lev, m1, m2);
  lts_classify(E, lev, m1, m2, cid, cls);
  const std::vector<int32_t> ord(E, c->P);
  dg_status s = ctx_relayout(c, ord, cid);
  if (s) return s;
  // each class is one internal range
  for (auto& k : cls) k.e0 = k.e1 = -1;
  for (int i = 0; i < E; i++) {
    LtsClass& k = cls[cid[c->perm[i]]];
    if (k.e0 < 0) k.e0 = i;
    else if (k.e1 != i) return fail(c, DG_ERR_ARG, "local time stepping: class %d not contiguous", cid[c->perm[i]]);
    k.e1 = i + 1;
  }
  std::vector<double> scale(E);
  for (int i = 0; i < E; i++) scale[i] = (double)(1 << lev[c->perm[i]]);
  LtsPlan plan;
  lts_plan(nlev, cls, plan);
  if ((s = dalloc(c, &c->dtscale, (size_t)E)) || (s = dalloc(c, &c->lts_tiles, plan.tiles.size())) ||
      (s = dalloc(c, &c->lts_ctiles, plan.ctiles.size() + 1)))
    return s;
  HIPCHECK(c, hipMemcpy(c->dtscale, scale.data(), (size_t)E * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(c->lts_tiles, plan.tiles.data(), plan.tiles.size() * sizeof(LtsTile), hipMemcpyHostToDevice));
  if (!plan.ctiles.empty())
    HIPCHECK(c, hipMemcpy(c->lts_ctiles, plan.ctiles.data(), plan.ctiles.size() * sizeof(LtsTile),
                          hipMemcpyHostToDevice));
  drop_graphs(c);  // captured local-time-stepping steps hold the old plan's tiles
  c->lts_plan = std::move(plan);
  c->lts_nlev = nlev;
  c->lts_level = lev;
  c->lts_cls = cls;
  if (levels_out) std::memcpy(levels_out, lev.data(), (size_t)E * 4);
  return DG_OK;
}

// one DG_RK_TVD2_LTS step of size h on the stream: per stage one launch over
// the stage's tiles (input Un, the span-start states, at even stages; A64, the
// predictors, at odd stages) and, at odd stages, one commit launch; acc holds
// each element's span increment
static int enqueue_lts_step(dg_ctx* c, double h) {
  const LtsPlan& P = c->lts_plan;
  StageArgs a = base_args(c);
  a.scheme = SCHEME_LTS;
  a.dt = h;
  for (int i = 0; i < P.S; i++) {
    const bool odd = i & 1;
    a.lts_tiles = c->lts_tiles + P.off[i];
    a.Uin = odd ? (const void*)c->A64 : (const void*)c->Un;
    a.Uout = odd ? (void*)c->Un : (void*)c->A64;
    if (launch_tri_stage_lts(a, c->law, P.off[i + 1] - P.off[i], c->stream)) return -1;
    if (launch_lts_commit(c->E, c->Nf * c->NP, c->lts_ctiles + P.coff[i], P.coff[i + 1] - P.coff[i], c->acc, c->Un,
                          c->stream))
      return -1;
  }
  return 0;
}

dg_status dg_select_orders(dg_ctx* c, const dg_sensor_opts* o, int apply, int32_t* orders_out) {
  if (!ready(c) || !o) return fail(c, DG_ERR_ARG, "dg_select_orders: bad argument");
  if (apply && c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "order changes run meshes without hanging nodes");
  if (c->dim != 2) return fail(c, DG_ERR_UNSUPPORTED, "built for triangles");
  if (o->pmin < 1 || o->pmax > MAXP_TRI || o->pmin > o->pmax) return fail(c, DG_ERR_ARG, "orders must lie in 1..%d", MAXP_TRI);
  if (apply && c->prec != DG_PREC_FP64) return fail(c, DG_ERR_UNSUPPORTED, "order changes run in the FP64 form only");
  if (apply && c->pvel) return fail(c, DG_ERR_UNSUPPORTED, "a velocity field runs uniform orders only");
  HIPCHECK(c, hipSetDevice(c->device));
  SensorArgs a{};
  a.E = c->E;
  for (int p = 0; p < 6; p++) {
    a.bstart[p] = c->bstart[p];
    a.boff[p] = c->boff[p];
    a.s_hi[p] = p >= 1 && p <= MAXP_TRI ? o->s_hi[p - 1] : 0.0;
    a.s_lo[p] = p >= 1 && p <= MAXP_TRI ? o->s_lo[p - 1] : 0.0;
  }
  a.pmin = o->pmin;
  a.pmax = o->pmax;
  a.U = c->Un;
  a.orders = c->sens;
  if (launch_sensor(a, c->stream)) return fail(c, DG_ERR_HIP, "sensor launch failed");
  std::vector<uint8_t> h(c->E);
  dg_status s = fetch_flags(c, c->sens, h.data());
  if (s) return s;
  std::vector<int32_t> neword(h.begin(), h.end());
  if (orders_out) std::memcpy(orders_out, neword.data(), neword.size() * 4);
  if (apply && neword != c->m_order) return ctx_rebucket(c, neword);
  return DG_OK;
}

// ---------------------------------------------------------------------------
// residual derivatives and the discrete adjoint of RK4 (no counterpart in the
// reference)

static void to_soa(const dg_ctx* c, const double* U, std::vector<double>& h) {
  h.resize(c->ndof);
  for_each_coeff(c, [&](int, long ci, long si) { h[si] = U[ci]; });
}

static dg_status upload_soa(dg_ctx* c, const double* U, double* dst) {
  std::vector<double> h;
  to_soa(c, U, h);
  HIPCHECK(c, hipMemcpyAsync(dst, h.data(), c->ndof * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return DG_OK;
}

static dg_status adj_buffers(dg_ctx* c) {
  if (c->lam) return DG_OK;
  dg_status s;
  const size_t nd = c->ndof;
  if ((s = dalloc(c, &c->lam, nd)) || (s = dalloc(c, &c->mu, nd)) || (s = dalloc(c, &c->za, nd)) ||
      (s = dalloc(c, &c->zb, nd)) || (s = dalloc(c, &c->wv, nd)))
    return s;
  return DG_OK;
}

static AdjArgs adj_args(dg_ctx* c) {
  AdjArgs a{};
  a.E = c->E;
  a.P = c->P;
  a.tab = c->tab;
  a.el.mg = c->mg;
  a.el.detj = c->detj;
  a.el.size = c->size;
  a.el.link = c->link;
  a.law = c->lp;
  a.mode = ADJ_PLAIN;
  return a;
}

dg_status dg_residual_jvp(dg_ctx* c, const double* U, const double* dU, double* dR) {
  if (!ready(c) || !U || !dU || !dR) return fail(c, DG_ERR_ARG, "dg_residual_jvp: bad argument");
  if (c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "residual derivatives run meshes without hanging nodes");
  if (c->mixed || c->dim != 2 || c->pvel)
    return fail(c, DG_ERR_UNSUPPORTED, "derivatives are built for uniform-order triangles (constant velocity)");
  HIPCHECK(c, hipSetDevice(c->device));
  dg_status s = adj_buffers(c);
  if (s || (s = upload_soa(c, U, c->lam)) || (s = upload_soa(c, dU, c->za))) return s;
  AdjArgs a = adj_args(c);
  a.U = c->lam;
  a.Z = c->za;
  a.W = c->wv;
  if (launch_tri_jvp(a, c->law, c->stream)) return fail(c, DG_ERR_HIP, "jvp launch failed");
  return fetch_soa(c, c->wv, dR);
}

dg_status dg_residual_vjp(dg_ctx* c, const double* U, const double* z, double* w) {
  if (!ready(c) || !U || !z || !w) return fail(c, DG_ERR_ARG, "dg_residual_vjp: bad argument");
  if (c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "residual derivatives run meshes without hanging nodes");
  if (c->mixed || c->dim != 2 || c->pvel)
    return fail(c, DG_ERR_UNSUPPORTED, "derivatives are built for uniform-order triangles (constant velocity)");
  HIPCHECK(c, hipSetDevice(c->device));
  dg_status s = adj_buffers(c);
  if (s || (s = upload_soa(c, U, c->lam)) || (s = upload_soa(c, z, c->za))) return s;
  AdjArgs a = adj_args(c);
  a.U = c->lam;
  a.Z = c->za;
  a.W = c->wv;
  if (launch_tri_vjp(a, c->law, c->adj_nd, c->stream)) return fail(c, DG_ERR_HIP, "vjp launch failed");
  return fetch_soa(c, c->wv, w);
}

// one forward RK4 step from seg slot (j,0) writing the stage inputs (j,1..3) and
// the next step's state to `next`
static int enqueue_rk4_stored(dg_ctx* c, double dt, double* s0, double* s1, double* s2, double* s3, double* next) {
  StageArgs a = base_args(c);
  a.dt = dt;
  a.scheme = DG_RK4;
  a.Un = s0;
  double* in[4] = {s0, s1, s2, s3};
  double* out[4] = {s1, s2, s3, next};
  for (int st = 0; st < 4; st++) {
    if (!out[st]) break;
    a.stage = st;
    a.Uin = in[st];
    a.Uout = out[st];
    a.resid_part = nullptr;
    if (launch_stage(c, a, 0)) return -1;
  }
  return 0;
}

// mass-weighted density misfit J = 1/2 sum_e |J_e| sum_k (rho_ek - rho*_ek)^2 and its
// gradient |J_e| (rho_ek - rho*_ek) (zero on the other fields), caller layout
static double density_misfit(const dg_ctx* c, const double* U, const double* target, double* grad) {
  const int BS = c->Nf * c->NP;
  double J = 0.0;
  std::fill(grad, grad + c->ndof, 0.0);
  for (int e = 0; e < c->E; e++) {
    const double* X = &c->xyz[(size_t)e * 6];
    const double ad = std::fabs((X[2] - X[0]) * (X[5] - X[1]) - (X[4] - X[0]) * (X[3] - X[1]));
    for (int k = 0; k < c->NP; k++) {
      const size_t i = (size_t)e * BS + k;
      const double d = U[i] - target[i];
      J += 0.5 * ad * d * d;
      grad[i] = ad * d;
    }
  }
  return J;
}

dg_status dg_adjoint_forward(dg_ctx* c, double dt, int nsteps, int ckpt_every, const double* target, double* J_out) {
  if (!ready(c) || !(dt > 0) || nsteps <= 0 || ckpt_every <= 0)
    return fail(c, DG_ERR_ARG, "dg_adjoint_forward: bad argument");
  if (c->hang_n) return fail(c, DG_ERR_UNSUPPORTED, "the adjoint runs meshes without hanging nodes");
  if (c->prec != DG_PREC_FP64) return fail(c, DG_ERR_UNSUPPORTED, "the adjoint is built for the FP64 form");
  if (c->mixed || c->lim_on || c->dim != 2 || c->pvel)
    return fail(c, DG_ERR_UNSUPPORTED, "the adjoint is built for the unlimited uniform-order triangle residual");
  HIPCHECK(c, hipSetDevice(c->device));
  c->adj_ready = false;
  dg_status s = adj_buffers(c);
  if (s) return s;
  const int nck = (nsteps + ckpt_every - 1) / ckpt_every;
  const size_t need_ck = c->ndof * (size_t)nck, need_seg = c->ndof * (size_t)4 * std::min(ckpt_every, nsteps);
  if (need_ck > c->adj_ck_cap) {
    if (c->ckpt) release(c, c->ckpt);
    c->ckpt = nullptr;
    if ((s = dalloc(c, &c->ckpt, need_ck))) return s;
    c->adj_ck_cap = need_ck;
  }
  if (need_seg > c->adj_seg_cap) {
    if (c->seg) release(c, c->seg);
    c->seg = nullptr;
    if ((s = dalloc(c, &c->seg, need_seg))) return s;
    c->adj_seg_cap = need_seg;
  }
  c->adj_nck = nck;
  c->adj_every = ckpt_every;
  c->adj_nsteps = nsteps;
  c->adj_dt = dt;
  hipEvent_t e0, e1;
  HIPCHECK(c, hipEventCreate(&e0));
  HIPCHECK(c, hipEventCreate(&e1));
  HIPCHECK(c, hipEventRecord(e0, c->stream));
  for (int st = 0; st < nsteps; st++) {
    if (st % ckpt_every == 0)
      HIPCHECK(c, hipMemcpyAsync(c->ckpt + (size_t)(st / ckpt_every) * c->ndof, c->Un, c->ndof * 8,
                                 hipMemcpyDeviceToDevice, c->stream));
    if (enqueue_step(c, DG_RK4, dt, false)) return fail(c, DG_ERR_HIP, "forward step failed");
  }
  HIPCHECK(c, hipEventRecord(e1, c->stream));
  HIPCHECK(c, hipEventSynchronize(e1));
  HIPCHECK(c, hipEventElapsedTime(&c->adj_fwd_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->adj_seed.clear();
  if (target) {
    std::vector<double> U(c->ndof);
    if ((s = dg_get_state(c, U.data(), U.size()))) return s;
    c->adj_seed.resize(c->ndof);
    const double J = density_misfit(c, U.data(), target, c->adj_seed.data());
    if (J_out) *J_out = J;
  } else if (J_out) {
    *J_out = 0.0;
  }
  c->adj_ready = true;
  return DG_OK;
}

dg_status dg_adjoint_backward(dg_ctx* c, const double* dJdUT, double* dJdU0) {
  if (!ready(c) || !dJdU0) return fail(c, DG_ERR_ARG, "dg_adjoint_backward: bad argument");
  if (!c->adj_ready) return fail(c, DG_ERR_ARG, "dg_adjoint_backward needs a preceding dg_adjoint_forward");
  if (!dJdUT && c->adj_seed.empty())
    return fail(c, DG_ERR_ARG, "dg_adjoint_backward: no seed (pass dJdUT or a target to the forward)");
  HIPCHECK(c, hipSetDevice(c->device));
  dg_status s = upload_soa(c, dJdUT ? dJdUT : c->adj_seed.data(), c->lam);
  if (s) return s;
  const size_t nd = c->ndof;
  const double dt = c->adj_dt;
  const int C = c->adj_every, N = c->adj_nsteps;
  hipEvent_t e0, e1;
  HIPCHECK(c, hipEventCreate(&e0));
  HIPCHECK(c, hipEventCreate(&e1));
  HIPCHECK(c, hipEventRecord(e0, c->stream));
  double* lam = c->lam;
  double* mu = c->mu;
  for (int ck = c->adj_nck - 1; ck >= 0; ck--) {
    const int m = std::min(N, ck * C + C) - ck * C;
    // recompute the segment's stage states from its checkpoint (the last step's
    // output is not needed)
    HIPCHECK(c, hipMemcpyAsync(c->seg, c->ckpt + (size_t)ck * nd, nd * 8, hipMemcpyDeviceToDevice, c->stream));
    for (int j = 0; j < m; j++) {
      double* b = c->seg + (size_t)j * 4 * nd;
      double* next = (j + 1 < m) ? c->seg + (size_t)(j + 1) * 4 * nd : nullptr;
      if (enqueue_rk4_stored(c, dt, b, b + nd, b + 2 * nd, b + 3 * nd, next))
        return fail(c, DG_ERR_HIP, "recompute failed");
    }
    // reverse sweep: z_3 = dt/|J| lam/6, then stages 3..0 of the fused adjoint kernel
    for (int j = m - 1; j >= 0; j--) {
      double* b = c->seg + (size_t)j * 4 * nd;
      if (launch_tri_adj_scale(c->E, c->Nf * c->NP, 1.0 / 6.0, dt, c->detj, lam, c->za, c->stream))
        return fail(c, DG_ERR_HIP, "adjoint seed failed");
      AdjArgs a = adj_args(c);
      a.mode = ADJ_RK4;
      a.dt = dt;
      a.lam = lam;
      a.mu = mu;
      double* zin = c->za;
      double* zout = c->zb;
      for (int st = 3; st >= 0; st--) {
        a.stage = st;
        a.U = b + (size_t)st * nd;
        a.Z = zin;
        a.znext = zout;
        if (launch_tri_vjp(a, c->law, c->adj_nd, c->stream)) return fail(c, DG_ERR_HIP, "adjoint stage failed");
        std::swap(zin, zout);
      }
      std::swap(lam, mu);
    }
  }
  HIPCHECK(c, hipEventRecord(e1, c->stream));
  HIPCHECK(c, hipEventSynchronize(e1));
  HIPCHECK(c, hipEventElapsedTime(&c->adj_bwd_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->lam = lam;
  c->mu = mu;
  return fetch_soa(c, lam, dJdU0);
}

dg_status dg_adjoint_timing(const dg_ctx* c, float* fwd_ms, float* bwd_ms) {
  if (!c || !fwd_ms || !bwd_ms) return DG_ERR_ARG;
  *fwd_ms = c->adj_fwd_ms;
  *bwd_ms = c->adj_bwd_ms;
  return DG_OK;
}

// ---------------------------------------------------------------------------
// launch schedules and the tuner

// triangle tile shapes: 0 (4 waves), 1 (8 waves), and 2 (4 waves with the
// contractions on the matrix cores) for the FP32 form of the Euler law
// tile shapes of the triangle stage: 4 and 8 waves, and the matrix-core shape of
// the FP32 Euler form (not built with the characteristic far field)
static int tri_shapes(const dg_ctx* c) {
  return c->law == DG_LAW_EULER2D && c->prec == DG_PREC_FP32 && !c->lp.char_bc ? 3 : 2;
}

static bool schedule_ok(const dg_ctx* c, const dg_schedule& s) {
  const bool tiled = c->dim == 3 || (!c->mixed && c->kernel == 0);
  const int nshape = c->dim == 3 ? TET_SHAPES : tri_shapes(c);
  // mixed orders: shape 1 = one launch per order bucket (no persistent grid)
  if (c->mixed) {
    if (s.shape < 0 || s.shape > 1 || (s.shape == 1 && (s.grid || !c->side[0]))) return false;
  } else if (s.shape < 0 || s.shape >= nshape || (s.shape && !tiled)) {
    return false;
  }
  if (s.grid < 0 || (s.grid && !c->mixed)) return false;
  if (s.grid && ((size_t)s.grid + 1 > c->part_cap || s.grid > c->nchunks)) return false;
  return true;
}

static void apply_schedule(dg_ctx* c, const dg_schedule& s) {
  if (c->shape != s.shape || c->mix_grid != s.grid) {
    drop_graphs(c);  // captured with the old launch
  }
  c->shape = s.shape;
  c->mix_grid = s.grid;
}

dg_status dg_get_schedule(const dg_ctx* c, dg_schedule* out) {
  if (!c || !out) return DG_ERR_ARG;
  *out = dg_schedule{};
  out->shape = c->shape;
  out->grid = c->mix_grid;
  return DG_OK;
}

dg_status dg_set_schedule(dg_ctx* c, const dg_schedule* s) {
  if (!ready(c) || !s) return fail(c, DG_ERR_ARG, "dg_set_schedule: bad argument");
  if (!schedule_ok(c, *s)) return fail(c, DG_ERR_ARG, "schedule (shape %d, grid %d) does not fit this context", s->shape, s->grid);
  apply_schedule(c, *s);
  return DG_OK;
}

// the candidates of this context: tile shapes of the tiled kernels, workgroup
// counts of the persistent mixed-order launch
static std::vector<dg_schedule> tune_candidates(const dg_ctx* c) {
  std::vector<dg_schedule> v;
  if (c->mixed) {
    // the default (one workgroup per chunk, at most 2048), multiples of the CU
    // count below it, and a half and a quarter of the full count
    const int full = std::min<int>(c->nchunks, (int)c->part_cap - 1), def = mixed_grid(c->nchunks);
    std::vector<int> grids = {0};
    for (int g : {256, 512, 1024, 2048, 4096, full / 4, full / 2, full})
      if (g >= 1 && g <= full && g != def && std::find(grids.begin(), grids.end(), g) == grids.end()) grids.push_back(g);
    for (int g : grids) {
      dg_schedule s{};
      s.grid = g;
      if (schedule_ok(c, s)) v.push_back(s);
    }
    dg_schedule sb{};
    sb.shape = 1;  // bucket by bucket
    if (schedule_ok(c, sb)) v.push_back(sb);
  } else {
    for (int sh = 0; sh < (c->dim == 3 ? TET_SHAPES : tri_shapes(c)); sh++) {
      dg_schedule s{};
      s.shape = sh;
      if (schedule_ok(c, s)) v.push_back(s);
    }
  }
  return v;
}

static void tune_key(const dg_ctx* c, char* key, size_t n) {
  hipDeviceProp_t prop{};
  const char* arch = hipGetDeviceProperties(&prop, c->device) == hipSuccess ? prop.gcnArchName : "unknown";
  char orders[64];
  if (c->mixed)
    snprintf(orders, sizeof(orders), "P1-4:%d,%d,%d,%d", c->bstart[2] - c->bstart[1], c->bstart[3] - c->bstart[2],
             c->bstart[4] - c->bstart[3], c->bstart[5] - c->bstart[4]);
  else
    snprintf(orders, sizeof(orders), "P%d", c->P);
  snprintf(key, n, "%s dim%d law%d flux%d %s prec%d E%d lim%d kernel%d", arch, c->dim, c->m_law.law, c->m_law.flux,
           orders, c->prec, c->E, c->lim_on, c->kernel);
  for (char* q = key; *q; q++)
    if (*q == '\t' || *q == '\n') *q = ' ';
}

// the last entry of the cache for key, if any
static bool tune_cache_lookup(const char* path, const char* key, dg_schedule* out) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char line[512];
  bool found = false;
  const size_t kl = strlen(key);
  while (fgets(line, sizeof(line), f)) {
    if (strncmp(line, key, kl) != 0 || line[kl] != '\t') continue;
    dg_schedule s{};
    if (sscanf(line + kl + 1, "%d %d", &s.shape, &s.grid) == 2) {
      *out = s;
      found = true;
    }
  }
  fclose(f);
  return found;
}

dg_status dg_autotune(dg_ctx* c, int scheme, double dt, int steps, const char* cache, int flags, dg_tune_report* rep) {
  if (!ready(c) || steps <= 0 || !(dt > 0) || !rep) return fail(c, DG_ERR_ARG, "dg_autotune: bad argument");
  dg_status st = check_scheme(c, scheme);
  if (st) return st;
  HIPCHECK(c, hipSetDevice(c->device));
  std::memset(rep, 0, sizeof(*rep));
  tune_key(c, rep->key, sizeof(rep->key));
  const std::vector<dg_schedule> cand = tune_candidates(c);
  if (cache && !(flags & DG_TUNE_FORCE)) {
    dg_schedule s{};
    if (tune_cache_lookup(cache, rep->key, &s) && schedule_ok(c, s)) {
      apply_schedule(c, s);
      rep->best = s;
      rep->from_cache = 1;
      return DG_OK;
    }
  }
  // the state (and the limiter flags) before tuning, restored after every run
  double* save = nullptr;
  uint8_t* fsave = nullptr;
  if ((st = dalloc(c, &save, c->ndof))) return st;
  if (c->flags && (st = dalloc(c, &fsave, (size_t)c->E))) {
    release(c, save);
    return st;
  }
  auto restore = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(c->Un, save, c->ndof * 8, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess && fsave) e = hipMemcpyAsync(c->flags, fsave, (size_t)c->E, hipMemcpyDeviceToDevice, c->stream);
    return e;
  };
  hipError_t he = hipMemcpyAsync(save, c->Un, c->ndof * 8, hipMemcpyDeviceToDevice, c->stream);
  if (he == hipSuccess && fsave) he = hipMemcpyAsync(fsave, c->flags, (size_t)c->E, hipMemcpyDeviceToDevice, c->stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (he == hipSuccess) he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  const dg_schedule before{c->shape, c->mix_grid, {0, 0, 0, 0, 0, 0}};
  const int ncand = (int)std::min<size_t>(cand.size(), DG_TUNE_MAX_CAND);
  std::vector<float> bestms(ncand, 0.f);
  // warm the device (clocks, caches) with the current schedule, untimed, so the
  // first candidate is not timed cold
  if (he == hipSuccess && (st = run_steps(c, scheme, dt, 4 * steps)) == DG_OK) he = restore();
  // rounds over the candidates in turn; per candidate one untimed step after the
  // switch (the graph of the new launch is built there), then the timed steps;
  // each candidate keeps its best round
  constexpr int ROUNDS = 3;
  for (int r = 0; r < ROUNDS && he == hipSuccess && st == DG_OK; r++)
    for (int i = 0; i < ncand && he == hipSuccess && st == DG_OK; i++) {
      apply_schedule(c, cand[i]);
      float ms = 0.f;
      if ((st = run_steps(c, scheme, dt, 1)) || (he = restore()) != hipSuccess ||
          (he = hipEventRecord(e0, c->stream)) != hipSuccess || (st = run_steps(c, scheme, dt, steps)) ||
          (he = hipEventRecord(e1, c->stream)) != hipSuccess || (he = hipEventSynchronize(e1)) != hipSuccess ||
          (he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess || (he = restore()) != hipSuccess)
        break;
      if (r == 0 || ms < bestms[i]) bestms[i] = ms;
    }
  int best = -1;
  if (he == hipSuccess && st == DG_OK)
    for (int i = 0; i < ncand; i++) {
      rep->cand[i] = cand[i];
      rep->ms_per_step[i] = bestms[i] / steps;
      if (best < 0 || rep->ms_per_step[i] < rep->ms_per_step[best]) best = i;
    }
  rep->ncand = best < 0 ? 0 : ncand;
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  release(c, save);
  if (fsave) release(c, fsave);
  if (st != DG_OK || he != hipSuccess || best < 0) {
    apply_schedule(c, before);
    if (st != DG_OK) return st;
    return fail(c, DG_ERR_HIP, "dg_autotune: %s", best < 0 ? "no candidate" : hipGetErrorString(he));
  }
  apply_schedule(c, rep->cand[best]);
  rep->best = rep->cand[best];
  if (cache) {
    FILE* f = fopen(cache, "a");
    if (!f) return fail(c, DG_ERR_ARG, "dg_autotune: cannot append to %s", cache);
    fprintf(f, "%s\t%d %d\n", rep->key, rep->best.shape, rep->best.grid);
    fclose(f);
  }
  return DG_OK;
}

dg_status dg_copy_probe(int device, size_t n_doubles, float* ms_out) {
  if (n_doubles < 4 || !ms_out) return DG_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return DG_ERR_HIP;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n_doubles * 8) != hipSuccess) return DG_ERR_OOM;
  if (hipMalloc(&b, n_doubles * 8) != hipSuccess) {
    (void)hipFree(a);
    return DG_ERR_OOM;
  }
  hipEvent_t e0, e1;
  dg_status st = DG_OK;
  if (hipMemset(a, 0, n_doubles * 8) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    st = DG_ERR_HIP;
  } else {
    (void)hipEventRecord(e0, nullptr);
    if (launch_copy_probe(a, b, (long)n_doubles, nullptr)) st = DG_ERR_HIP;
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_out = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return st;
}

// ---------------------------------------------------------------------------
static uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unit_rand(uint64_t x) { return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0); }

dg_status dg_mesh_tri_structured(int nx, int ny, double x0, double y0, double lx, double ly, int periodic_x,
                                 int periodic_y, double jitter, uint64_t seed, const int32_t side_tags[4],
                                 int32_t* vids, double* xyz, int32_t* bc) {
  if (nx < 1 || ny < 1 || !(lx > 0) || !(ly > 0) || !vids || !xyz || jitter < 0 || jitter >= 0.25) return DG_ERR_ARG;
  if ((periodic_x && nx < 3) || (periodic_y && ny < 3)) return DG_ERR_ARG;
  const double hx = lx / nx, hy = ly / ny;
  auto lattice = [&](int i, int j, double* p) {
    p[0] = x0 + i * hx;
    p[1] = y0 + j * hy;
    if (i > 0 && i < nx && j > 0 && j < ny && jitter > 0) {
      uint64_t vid = (uint64_t)j * (nx + 1) + i;
      p[0] += (2.0 * unit_rand(seed ^ (2 * vid)) - 1.0) * jitter * hx;
      p[1] += (2.0 * unit_rand(seed ^ (2 * vid + 1)) - 1.0) * jitter * hy;
    }
  };
  auto topo = [&](int i, int j) {
    if (periodic_x && i == nx) i = 0;
    if (periodic_y && j == ny) j = 0;
    return (int32_t)(j * (nx + 1) + i);
  };
  // side of a lattice edge: 0 bottom, 1 right, 2 top, 3 left, -1 interior
  auto side = [&](int i0, int j0, int i1, int j1) {
    if (j0 == 0 && j1 == 0) return 0;
    if (i0 == nx && i1 == nx) return 1;
    if (j0 == ny && j1 == ny) return 2;
    if (i0 == 0 && i1 == 0) return 3;
    return -1;
  };
  size_t e = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int tri[2][3][2] = {{{i, j}, {i + 1, j}, {i + 1, j + 1}}, {{i, j}, {i + 1, j + 1}, {i, j + 1}}};
      for (int t = 0; t < 2; t++, e++) {
        for (int v = 0; v < 3; v++) {
          vids[e * 3 + v] = topo(tri[t][v][0], tri[t][v][1]);
          lattice(tri[t][v][0], tri[t][v][1], xyz + e * 6 + v * 2);
        }
        if (bc)
          for (int l = 0; l < 3; l++) {
            int sd = side(tri[t][l][0], tri[t][l][1], tri[t][(l + 1) % 3][0], tri[t][(l + 1) % 3][1]);
            bool per = (sd == 0 || sd == 2) ? periodic_y : (sd == 1 || sd == 3) ? periodic_x : true;
            bc[e * 3 + l] = (sd >= 0 && !per) ? (side_tags ? side_tags[sd] : DG_TAG_FARFIELD) : 0;
          }
      }
    }
  return DG_OK;
}

// Kuhn subdivision of an nx x ny x nz lattice of boxes on [0,lx]x[0,ly]x[0,lz]:
// every box is cut into the 6 tetrahedra that share its main diagonal
// (v000 -> v000+e_a -> v000+e_a+e_b -> v111 for the permutations (a,b,c)), so the
// faces of neighbouring boxes match.  Vertex ids are lattice ids.
dg_status dg_mesh_tet_structured(int nx, int ny, int nz, double lx, double ly, double lz, int32_t* vids,
                                 double* xyz) {
  if (nx < 1 || ny < 1 || nz < 1 || !(lx > 0) || !(ly > 0) || !(lz > 0) || !vids || !xyz) return DG_ERR_ARG;
  const double h[3] = {lx / nx, ly / ny, lz / nz};
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  size_t e = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int t = 0; t < 6; t++, e++) {
          int c[3] = {i, j, k};
          for (int v = 0; v < 4; v++) {
            if (v > 0 && v < 3) c[perm[t][v - 1]] += 1;
            if (v == 3) c[perm[t][2]] += 1;
            vids[e * 4 + v] = (int32_t)(((long)c[2] * (ny + 1) + c[1]) * (nx + 1) + c[0]);
            for (int d = 0; d < 3; d++) xyz[e * 12 + v * 3 + d] = c[d] * h[d];
          }
        }
  return DG_OK;
}

}  // extern "C"
