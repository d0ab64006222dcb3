// Checkpoint / restart and state export.
//
//  - dg_write_restart / dg_read_restart: the data stream of DGAnalysis::write and
//    DGAnalysis::read (DGPref.cc:712-768) with the cell records of
//    mMesh::writeMEntity + DGCell::write (mMesh.cc:401-412, DGCell.cc:1523-1540).
//  - dg_mesh_write_gmsh: the mesh stream (mImportExport::exportDGFile,
//    mImportExport.cc:611-642) as a Gmsh version-1 file that dg_mesh_read_gmsh
//    reads back, periodic sides included.
//  - dg_sample_lattice / dg_export_pos: the fields at the corners of the plot
//    sub-simplices (a device kernel) and the Gmsh post-processing views of
//    DGAnalysis::exportGmsh / exportGmsh_Volume (DGSensor.cc:337-525, 800-).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/dg_capi.h"
#include "dg_ctx.h"
#include "dg_fields.h"
#include "dg_sample.h"
#include "dg_tables.h"

namespace {

// entity types of mEntity::mType (mEntity.h:106)
enum { T_EDGE = 1, T_TRI = 2, T_TET = 4 };

// local face l of a tetrahedron: the face without vertex (l + 3) % 4; of a
// triangle: edge (v_l, v_{l+1})
const int TETF[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};

int face_vertices(int dim, int l, int* v) {
  if (dim == 2) {
    v[0] = l;
    v[1] = (l + 1) % 3;
    return 2;
  }
  for (int i = 0; i < 3; i++) v[i] = TETF[l][i];
  return 3;
}

std::array<int, 4> sorted_key(const int32_t* ids, int n) {
  std::array<int, 4> k{-1, -1, -1, -1};
  for (int i = 0; i < n; i++) k[i] = ids[i];
  std::sort(k.begin(), k.begin() + n);
  return k;
}

int np_of(int dim, int p) { return dim == 2 ? np_tri(p) : np_tet(p); }

// lattice point n of dim-simplex split nsub times (order: k outer, j, i inner)
void lattice_point(int dim, int nsub, int n, double* xi) {
  int c = 0;
  for (int k = 0; k <= (dim == 3 ? nsub : 0); k++)
    for (int j = 0; j <= nsub - k; j++)
      for (int i = 0; i <= nsub - k - j; i++, c++)
        if (c == n) {
          xi[0] = (double)i / nsub;
          xi[1] = (double)j / nsub;
          xi[2] = (double)k / nsub;
          return;
        }
}

int lattice_index(int dim, int nsub, int i, int j, int k) {
  int c = 0;
  for (int kk = 0; kk <= (dim == 3 ? nsub : 0); kk++)
    for (int jj = 0; jj <= nsub - kk; jj++)
      for (int ii = 0; ii <= nsub - kk - jj; ii++, c++)
        if (ii == i && jj == j && kk == k) return c;
  return -1;
}

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

void put_err(char* out, size_t len, const std::string& msg) {
  if (out && len) std::snprintf(out, len, "%s", msg.c_str());
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// checkpoint / restart

dg_status dg_write_restart(dg_ctx* c, const char* path, double t, int iter) {
  if (!ready(c) || !path) return fail(c, DG_ERR_ARG, "dg_write_restart: bad argument");
  std::vector<double> U(c->ndof);
  dg_status s = dg_get_state(c, U.data(), U.size());
  if (s) return s;
  File out(path, "w");
  if (!out.f) return fail(c, DG_ERR_ARG, "cannot write %s", path);
  const int nv = c->dim + 1, NF = (int)c->faces.size() / 4;
  // 17 significant digits: every coefficient reads back to the same double
  std::fprintf(out.f, "%.17g\n%d\n%d %d %d\n", t, iter, c->dim, c->E, NF);
  for (int e = 0; e < c->E; e++) {
    std::fprintf(out.f, "%d ", c->dim == 2 ? T_TRI : T_TET);
    // vertex ids at the file boundary are node numbers, id + 1 (Vertex::getId() of a
    // Gmsh-read reference mesh is the node number, mMesh.cc:401-412)
    for (int v = 0; v < nv; v++) std::fprintf(out.f, "%d ", c->m_vids[(size_t)e * nv + v] + 1);
    const int p = c->m_order[e];
    std::fprintf(out.f, "\n%d ", p);
    for (long i = c->coff[e]; i < c->coff[e + 1]; i++) std::fprintf(out.f, "%.17g ", U[i]);
    std::fputc('\n', out.f);
  }
  for (int i = 0; i < NF; i++) {
    const int e = c->faces[4 * i], l = c->faces[4 * i + 1];
    int fv[3];
    const int n = face_vertices(c->dim, l, fv);
    std::fprintf(out.f, "%d ", c->dim == 2 ? T_EDGE : T_TRI);
    for (int k = 0; k < n; k++) std::fprintf(out.f, "%d ", c->m_vids[(size_t)e * nv + fv[k]] + 1);
    std::fputc('\n', out.f);
  }
  if (std::ferror(out.f)) return fail(c, DG_ERR_ARG, "write error on %s", path);
  return DG_OK;
}

dg_status dg_read_restart(dg_ctx* c, const char* path, double* t_out, int* iter_out) {
  if (!ready(c) || !path) return fail(c, DG_ERR_ARG, "dg_read_restart: bad argument");
  File in(path, "r");
  if (!in.f) return fail(c, DG_ERR_ARG, "cannot open %s", path);
  double t = 0.0;
  int iter = 0, dim = 0, ne = 0, nf = 0;
  if (std::fscanf(in.f, "%lf %d %d %d %d", &t, &iter, &dim, &ne, &nf) != 5)
    return fail(c, DG_ERR_ARG, "%s: bad header", path);
  if (dim != c->dim || ne != c->E)
    return fail(c, DG_ERR_ARG, "%s: %dD mesh of %d elements, the context has a %dD mesh of %d", path, dim, ne, c->dim,
                c->E);
  const int nv = c->dim + 1;
  std::map<std::array<int, 4>, int> elem_of;
  for (int e = 0; e < c->E; e++) elem_of[sorted_key(&c->m_vids[(size_t)e * nv], nv)] = e;
  std::vector<int32_t> ord(c->E, 0);
  std::vector<std::vector<double>> coef(c->E);
  const int maxp = c->dim == 2 ? MAXP_TRI : 3;
  for (int r = 0; r < ne; r++) {
    int type = 0, p = 0;
    int32_t ids[4];
    if (std::fscanf(in.f, "%d", &type) != 1 || type != (dim == 2 ? T_TRI : T_TET))
      return fail(c, DG_ERR_ARG, "%s: cell record %d: bad entity type", path, r);
    for (int v = 0; v < nv; v++) {
      if (std::fscanf(in.f, "%d", &ids[v]) != 1) return fail(c, DG_ERR_ARG, "%s: cell record %d truncated", path, r);
      ids[v] -= 1;  // node number -> vertex id
    }
    auto it = elem_of.find(sorted_key(ids, nv));
    if (it == elem_of.end()) return fail(c, DG_ERR_ARG, "%s: cell record %d: no element with these vertices", path, r);
    const int e = it->second;
    if (ord[e]) return fail(c, DG_ERR_ARG, "%s: element %d appears twice", path, e);
    if (!std::equal(ids, ids + nv, &c->m_vids[(size_t)e * nv]))
      return fail(c, DG_ERR_ARG, "%s: cell record %d: vertex order differs from element %d", path, r, e);
    if (std::fscanf(in.f, "%d", &p) != 1 || p < 1 || p > maxp)
      return fail(c, DG_ERR_ARG, "%s: cell record %d: bad order", path, r);
    ord[e] = p;
    coef[e].resize((size_t)c->Nf * np_of(dim, p));
    for (double& v : coef[e])
      if (std::fscanf(in.f, "%lf", &v) != 1) return fail(c, DG_ERR_ARG, "%s: cell record %d truncated", path, r);
  }
  // the face records that follow are the mesh's, which the context already has
  if (ord != c->m_order) {
    if (c->pvel) return fail(c, DG_ERR_UNSUPPORTED, "%s: orders differ and a velocity field runs one order", path);
    if (c->dim != 2) return fail(c, DG_ERR_UNSUPPORTED, "%s: orders differ from the context's (tetrahedra: one order)", path);
    if (c->prec != DG_PREC_FP64) return fail(c, DG_ERR_UNSUPPORTED, "%s: order changes run in the FP64 form only", path);
    dg_status s = ctx_rebucket(c, ord);
    if (s) return s;
  }
  std::vector<double> U(c->ndof);
  for (int e = 0; e < c->E; e++) std::copy(coef[e].begin(), coef[e].end(), U.begin() + c->coff[e]);
  dg_status s = dg_set_state(c, U.data(), U.size());
  if (s) return s;
  if (t_out) *t_out = t;
  if (iter_out) *iter_out = iter;
  return DG_OK;
}

// ---------------------------------------------------------------------------
// mesh file

dg_status dg_mesh_write_gmsh(const dg_mesh_desc* m, const char* path, char* err_out, size_t err_len) {
  if (!m || !path || !m->vids || !m->xyz || m->n_elems <= 0 || (m->dim != 2 && m->dim != 3)) {
    put_err(err_out, err_len, "bad argument");
    return DG_ERR_ARG;
  }
  const int dim = m->dim, nv = dim + 1, E = m->n_elems;
  // one node per (vertex id, coordinates).  The first position met of vertex id
  // v is node v + 1; periodic images at other positions get numbers above every
  // v + 1 and are joined again by the 510 / 610 tags below, so the reader (whose
  // vertex id is the smallest node number of a class, minus one) returns the
  // same vertex ids.
  int32_t vmax = -1;
  for (size_t i = 0; i < (size_t)E * nv; i++) {
    if (m->vids[i] < 0 || m->vids[i] == INT32_MAX) {
      put_err(err_out, err_len, "vertex ids must lie in 0..2^31-2");
      return DG_ERR_ARG;
    }
    vmax = std::max(vmax, m->vids[i]);
  }
  std::map<std::tuple<int, double, double, double>, long> node_of;
  std::map<int, bool> seen;
  long next_extra = (long)vmax + 2;
  std::vector<long> enode((size_t)E * nv);
  for (int e = 0; e < E; e++)
    for (int v = 0; v < nv; v++) {
      const int vid = m->vids[(size_t)e * nv + v];
      const double* x = m->xyz + ((size_t)e * nv + v) * dim;
      auto key = std::make_tuple(vid, x[0], x[1], dim == 3 ? x[2] : 0.0);
      auto it = node_of.find(key);
      if (it == node_of.end()) {
        const long id = seen.count(vid) ? next_extra++ : (long)vid + 1;
        seen[vid] = true;
        it = node_of.emplace(key, id).first;
      }
      enode[(size_t)e * nv + v] = it->second;
    }
  // faces by sorted vertex ids
  std::map<std::array<int, 4>, std::vector<std::pair<int, int>>> sides;
  for (int e = 0; e < E; e++)
    for (int l = 0; l < nv; l++) {
      int fv[3];
      int32_t ids[3];
      const int n = face_vertices(dim, l, fv);
      for (int k = 0; k < n; k++) ids[k] = m->vids[(size_t)e * nv + fv[k]];
      sides[sorted_key(ids, n)].push_back({e, l});
    }
  struct Rec {
    int type, tag;
    std::vector<long> nodes;
  };
  std::vector<Rec> recs;
  for (auto& kv : sides) {
    const auto& s = kv.second;
    if (s.size() > 2) {
      put_err(err_out, err_len, "a face is shared by more than two elements");
      return DG_ERR_ARG;
    }
    auto face_nodes = [&](int e, int l) {
      int fv[3];
      const int n = face_vertices(dim, l, fv);
      std::vector<long> r;
      for (int k = 0; k < n; k++) r.push_back(enode[(size_t)e * nv + fv[k]]);
      return r;
    };
    if (s.size() == 1) {
      const int tag = m->bc ? m->bc[(size_t)s[0].first * nv + s[0].second] : DG_TAG_FARFIELD;
      recs.push_back({dim == 2 ? 1 : 2, tag, face_nodes(s[0].first, s[0].second)});
      continue;
    }
    std::vector<long> a = face_nodes(s[0].first, s[0].second), b = face_nodes(s[1].first, s[1].second);
    std::vector<long> as = a, bs = b;
    std::sort(as.begin(), as.end());
    std::sort(bs.begin(), bs.end());
    if (as == bs) continue;  // an ordinary interior face
    if (dim == 3) {
      put_err(err_out, err_len, "periodic tetrahedral meshes have no Gmsh tags here");
      return DG_ERR_UNSUPPORTED;
    }
    // a periodic pair: the translation between the two images of one vertex id
    const int e0 = s[0].first, e1 = s[1].first;
    double d[2][2];
    for (int q = 0; q < 2; q++) {
      const int vid = m->vids[(size_t)e0 * 3 + (s[0].second + q) % 3];
      int v1 = -1;
      for (int v = 0; v < 3; v++)
        if (m->vids[(size_t)e1 * 3 + v] == vid) v1 = v;
      const double* x0 = m->xyz + ((size_t)e0 * 3 + (s[0].second + q) % 3) * 2;
      const double* x1 = m->xyz + ((size_t)e1 * 3 + v1) * 2;
      d[q][0] = x1[0] - x0[0];
      d[q][1] = x1[1] - x0[1];
    }
    const double tol = 1e-10;
    const bool same = std::fabs(d[0][0] - d[1][0]) < tol && std::fabs(d[0][1] - d[1][1]) < tol;
    const bool in_x = std::fabs(d[0][1]) < tol && std::fabs(d[0][0]) >= tol;
    const bool in_y = std::fabs(d[0][0]) < tol && std::fabs(d[0][1]) >= tol;
    if (!same || (!in_x && !in_y)) {
      put_err(err_out, err_len, "a periodic face pair is not a translation along x or y");
      return DG_ERR_UNSUPPORTED;
    }
    // sides translated along x pair by equal y (510), along y by equal x (610)
    const int tag = in_x ? 510 : 610;
    recs.push_back({1, tag, a});
    recs.push_back({1, tag, b});
  }
  File out(path, "w");
  if (!out.f) {
    put_err(err_out, err_len, std::string("cannot write ") + path);
    return DG_ERR_ARG;
  }
  std::vector<std::pair<long, const std::tuple<int, double, double, double>*>> nodes;
  for (auto& kv : node_of) nodes.push_back({kv.second, &kv.first});
  std::sort(nodes.begin(), nodes.end());
  std::fprintf(out.f, "$NOD\n%zu\n", nodes.size());
  for (const auto& nd : nodes)
    std::fprintf(out.f, "%ld %.17g %.17g %.17g\n", nd.first, std::get<1>(*nd.second), std::get<2>(*nd.second),
                 std::get<3>(*nd.second));
  std::fprintf(out.f, "$ENDNOD\n$ELM\n%zu\n", recs.size() + (size_t)E);
  size_t id = 1;
  for (const Rec& r : recs) {
    std::fprintf(out.f, "%zu %d %d 1 %zu", id++, r.type, r.tag, r.nodes.size());
    for (long n : r.nodes) std::fprintf(out.f, " %ld", n);
    std::fputc('\n', out.f);
  }
  for (int e = 0; e < E; e++) {
    std::fprintf(out.f, "%zu %d 100 1 %d", id++, dim == 2 ? 2 : 4, nv);
    for (int v = 0; v < nv; v++) std::fprintf(out.f, " %ld", enode[(size_t)e * nv + v]);
    std::fputc('\n', out.f);
  }
  std::fprintf(out.f, "$ENDELM\n");
  if (std::ferror(out.f)) {
    put_err(err_out, err_len, std::string("write error on ") + path);
    return DG_ERR_ARG;
  }
  return DG_OK;
}

// ---------------------------------------------------------------------------
// state export

dg_status dg_norms(dg_ctx* c, double t, int exact_field_id, const double* exact_par, double* l1, double* l2,
                   double* linf, double* totals, double* energy) {
  if (!ready(c)) return fail(c, DG_ERR_ARG, "dg_norms: bad context");
  if (totals || energy) {
    TotalsArgs a{};
    a.E = c->E;
    a.NF = c->Nf;
    for (int p = 0; p < 6; p++) {
      a.bstart[p] = c->bstart[p];
      a.boff[p] = c->boff[p];
      a.np[p] = p >= 1 && p <= 4 ? np_of(c->dim, p) : 0;
    }
    // int phi_0 over the reference simplex of measure 1/2 or 1/6, phi_0 = 1/sqrt(measure)
    a.phi0_int = c->dim == 2 ? std::sqrt(0.5) : std::sqrt(1.0 / 6.0);
    a.U = c->Un;
    a.detj = c->detj;
    const int nb = (c->E + TOTALS_BLOCK - 1) / TOTALS_BLOCK;
    const size_t n = (size_t)nb * 2 * c->Nf;
    HIPCHECK(c, hipSetDevice(c->device));
    double* dpart = nullptr;
    if (hipMalloc(&dpart, n * 8) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, DG_ERR_OOM, "dg_norms: out of device memory");
    }
    a.part = dpart;
    std::vector<double> h(n);
    hipError_t e = launch_totals(a, c->stream) ? hipErrorLaunchFailure : hipSuccess;
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dpart, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dpart);
    if (e != hipSuccess) return fail(c, DG_ERR_HIP, "dg_norms: %s", hipGetErrorString(e));
    for (int j = 0; j < 2 * c->Nf; j++) {
      double acc = 0.0;
      for (int b = 0; b < nb; b++) acc += h[(size_t)b * 2 * c->Nf + j];
      if (j < c->Nf && totals) totals[j] = acc;
      if (j >= c->Nf && energy) energy[j - c->Nf] = acc;
    }
  }
  if (l1 || l2 || linf) {
    if (exact_field_id < 0) return fail(c, DG_ERR_ARG, "dg_norms: error norms need an exact field");
    return ctx_error_norms(c, exact_field_id, exact_par, t, l1, l2, linf);
  }
  return DG_OK;
}

size_t dg_lattice_points(int dim, int nsub) {
  if ((dim != 2 && dim != 3) || nsub < 1 || nsub > SAMPLE_MAX_NSUB) return 0;
  return (size_t)lattice_points(dim, nsub);
}

dg_status dg_sample_lattice(dg_ctx* c, int nsub, double* out, size_t n) {
  if (!ready(c) || !out) return fail(c, DG_ERR_ARG, "dg_sample_lattice: bad argument");
  if (nsub < 1 || nsub > SAMPLE_MAX_NSUB) return fail(c, DG_ERR_ARG, "nsub must lie in 1..%d", SAMPLE_MAX_NSUB);
  const int NN = lattice_points(c->dim, nsub);
  const size_t total = (size_t)c->E * c->Nf * NN;
  if (n != total) return fail(c, DG_ERR_ARG, "dg_sample_lattice: need %zu doubles", total);
  SampleArgs a{};
  a.E = c->E;
  a.NF = c->Nf;
  a.NN = NN;
  // basis of every order present at the lattice points
  std::vector<double> tab;
  for (int p = 0; p < 6; p++) {
    a.bstart[p] = c->bstart[p];
    a.boff[p] = c->boff[p];
    if (p < 1 || p > 4 || c->bstart[p + 1] == c->bstart[p]) continue;
    const int np = np_of(c->dim, p);
    a.np[p] = np;
    a.toff[p] = (int)tab.size();
    tab.resize(tab.size() + (size_t)NN * np);
    for (int q = 0; q < NN; q++) {
      double xi[3];
      lattice_point(c->dim, nsub, q, xi);
      basis(c->dim, p, xi, &tab[a.toff[p] + (size_t)q * np], nullptr);
    }
  }
  HIPCHECK(c, hipSetDevice(c->device));
  double *dtab = nullptr, *dout = nullptr;
  int32_t* dperm = nullptr;
  auto cleanup = [&]() {
    if (dtab) (void)hipFree(dtab);
    if (dout) (void)hipFree(dout);
    if (dperm) (void)hipFree(dperm);
  };
  hipError_t e = hipMalloc(&dtab, std::max<size_t>(tab.size(), 1) * 8);
  if (e == hipSuccess) e = hipMalloc(&dout, total * 8);
  if (e == hipSuccess) e = hipMalloc(&dperm, (size_t)c->E * 4);
  if (e != hipSuccess) {
    cleanup();
    (void)hipGetLastError();
    return fail(c, DG_ERR_OOM, "dg_sample_lattice: out of device memory");
  }
  e = hipMemcpyAsync(dtab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dperm, c->perm.data(), (size_t)c->E * 4, hipMemcpyHostToDevice, c->stream);
  a.U = c->Un;
  a.tab = dtab;
  a.perm = dperm;
  a.out = dout;
  if (e == hipSuccess && launch_sample(a, c->stream)) e = hipErrorLaunchFailure;
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, total * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  cleanup();
  if (e != hipSuccess) return fail(c, DG_ERR_HIP, "dg_sample_lattice: %s", hipGetErrorString(e));
  return DG_OK;
}

dg_status dg_export_pos(dg_ctx* c, const char* prefix, int iteration, int nsub, double t, int exact_field_id,
                        const double* exact_par) {
  if (!ready(c) || !prefix) return fail(c, DG_ERR_ARG, "dg_export_pos: bad argument");
  const int dim = c->dim, nv = dim + 1, E = c->E, NF = c->Nf;
  int pmax = 1;
  for (int p : c->m_order) pmax = std::max(pmax, p);
  // sub-simplices per element: the element order in 2D (DGSensor.cc:362-366),
  // 1 or 8 tetrahedra in 3D (DGSensor.cc:816-817)
  if (nsub <= 0) nsub = dim == 2 ? pmax : (pmax <= 1 ? 1 : 2);
  if (nsub > SAMPLE_MAX_NSUB || (dim == 3 && nsub > 2))
    return fail(c, DG_ERR_ARG, "dg_export_pos: nsub must lie in 1..%d (tetrahedra: 1 or 2)", dim == 2 ? SAMPLE_MAX_NSUB : 2);
  const int NN = lattice_points(dim, nsub);
  std::vector<double> val((size_t)E * NF * NN);
  dg_status s = dg_sample_lattice(c, nsub, val.data(), val.size());
  if (s) return s;
  std::vector<uint8_t> flags(E, 0);
  if (dim == 2) {
    s = dg_limiter_flags(c, flags.data());
    if (s) return s;
  }
  // physical coordinates of the lattice points
  std::vector<double> xq((size_t)E * NN * 3);
  for (int e = 0; e < E; e++) {
    const double* X = &c->xyz[(size_t)e * nv * dim];
    for (int q = 0; q < NN; q++) {
      double xi[3];
      lattice_point(dim, nsub, q, xi);
      for (int d = 0; d < 3; d++) {
        double x = d < dim ? X[d] : 0.0;
        for (int v = 1; v < nv && d < dim; v++) x += xi[v - 1] * (X[v * dim + d] - X[d]);
        xq[((size_t)e * NN + q) * 3 + d] = x;
      }
    }
  }
  // plot sub-simplices as lattice indices
  std::vector<std::array<int, 4>> sub;
  if (dim == 2) {
    for (int j = 0; j < nsub; j++)
      for (int i = 0; i + j < nsub; i++) {
        sub.push_back({lattice_index(2, nsub, i, j, 0), lattice_index(2, nsub, i + 1, j, 0),
                       lattice_index(2, nsub, i, j + 1, 0), -1});
        if (i + j + 1 < nsub)
          sub.push_back({lattice_index(2, nsub, i + 1, j, 0), lattice_index(2, nsub, i + 1, j + 1, 0),
                         lattice_index(2, nsub, i, j + 1, 0), -1});
      }
  } else if (nsub == 1) {
    sub.push_back({0, 1, 2, 3});
  } else {
    // red refinement: four corner tetrahedra and the octahedron cut along its
    // (0,1,0)-(1,0,1) diagonal
    auto L = [&](int i, int j, int k) { return lattice_index(3, 2, i, j, k); };
    const int a = L(0, 0, 0), b = L(2, 0, 0), cc = L(0, 2, 0), d = L(0, 0, 2);
    const int ab = L(1, 0, 0), ac = L(0, 1, 0), ad = L(0, 0, 1), bc = L(1, 1, 0), bd = L(1, 0, 1), cd = L(0, 1, 1);
    sub = {{a, ab, ac, ad}, {ab, b, bc, bd}, {ac, bc, cc, cd}, {ad, bd, cd, d},
           {ac, bd, ab, bc}, {ac, bd, bc, cd}, {ac, bd, cd, ad}, {ac, bd, ad, ab}};
  }
  // views: (name, components); the per-point function fills comps values
  const LawParams& lp = c->lp;
  struct View {
    const char* name;
    int ncomp;
  };
  std::vector<View> views;
  const bool have_exact = exact_field_id >= 0;
  if (c->law == DG_LAW_EULER2D) {
    // Euler::getNameAndSize / Euler2d::getIthFieldOfInterest (EulerLaw.cc:1240-1340)
    views = {{"velocity", 3}, {"pressure", 1}, {"mach", 1}, {"density", 1}, {"error", 1}, {"entropy", 1}, {"limit", 1}};
  } else if (c->law == DG_LAW_SCALAR) {
    // ScalarLaw::getNameAndSize / getIthFieldOfInterest (ScalarLaw.cc:111-165)
    if (have_exact) views = {{"error", 1}, {"exact", 1}};
    views.push_back({"numEr", 1});
    views.push_back({"function", 1});
    views.push_back({"limit", 1});
  } else if (c->law == DG_LAW_MAXWELL2D) {
    // Maxwell::getNameAndSize / Maxwell2d::getIthFieldOfInterest
    // (MaxwellLaw.cc:256-330): Ez^2, Hx, Hy, Ez (its fifth view is the order)
    views = {{"Intensity_RefCond", 1}, {"HxField_RefCond", 1}, {"HyField_RefCond", 1}, {"EField_RefCond", 1}};
  } else if (c->law == DG_LAW_BURGERS2D) {
    // Burgers2D::getNameAndSize / getIthFieldOfInterest (Burgers2D.cc:121-200)
    views = {{"function", 1}};
    if (have_exact) {
      views.push_back({"error", 1});
      views.push_back({"exact", 1});
    }
    views.push_back({"numEr", 1});
  } else {
    // 3D Maxwell: the field vectors and the energy density (no 3D counterpart
    // in the reference, whose 2D list is MaxwellLaw.cc:256-330)
    views = {{"E", 3}, {"H", 3}, {"energy", 1}};
    if (have_exact) views.push_back({"error", 1});
  }
  views.push_back({"order", 1});
  double exact[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (!exact_par) exact_par = zero;
  auto point_value = [&](int vi, int e, int q, double* r) -> int {
    double u[8];
    for (int f = 0; f < NF; f++) u[f] = val[((size_t)e * NF + f) * NN + q];
    const std::string nm = views[vi].name;
    const double* x = &xq[((size_t)e * NN + q) * 3];
    if (nm == "order") return r[0] = c->m_order[e], 0;
    if (nm == "limit") return r[0] = flags[e], 0;
    if (have_exact && (nm == "error" || nm == "exact") && c->law != DG_LAW_EULER2D)
      if (dgp::exact_field(exact_field_id, exact_par, lp, x, t, exact)) return -1;
    if (c->law == DG_LAW_EULER2D) {
      const double g = lp.gamma, *ff = lp.farfield;
      const double inv = 1.0 / u[0], ke = 0.5 * (u[1] * u[1] + u[2] * u[2]) * inv;
      const double p = (g - 1.0) * (u[3] - ke);
      if (nm == "velocity") r[0] = u[1] * inv, r[1] = u[2] * inv, r[2] = 0.0;
      else if (nm == "pressure") r[0] = p;
      else if (nm == "mach") r[0] = std::sqrt(2.0 * ke * inv / (g * p * inv));
      else if (nm == "density") r[0] = u[0];
      else if (nm == "error") r[0] = u[1] - ff[1];
      else if (nm == "entropy") {
        const double pff = (g - 1.0) * (ff[3] - 0.5 * (ff[1] * ff[1] + ff[2] * ff[2]) / ff[0]);
        r[0] = p / pff / std::pow(u[0] / ff[0], g) - 1.0;
      }
    } else if (c->law == DG_LAW_SCALAR || c->law == DG_LAW_BURGERS2D) {
      if (nm == "error") r[0] = exact[0] - u[0];
      else if (nm == "exact") r[0] = exact[0];
      else r[0] = u[0];
    } else if (c->law == DG_LAW_MAXWELL2D) {
      if (nm == "Intensity_RefCond") r[0] = u[2] * u[2];
      else if (nm == "HxField_RefCond") r[0] = u[0];
      else if (nm == "HyField_RefCond") r[0] = u[1];
      else r[0] = u[2];
    } else {
      if (nm == "E") r[0] = u[0], r[1] = u[1], r[2] = u[2];
      else if (nm == "H") r[0] = u[3], r[1] = u[4], r[2] = u[5];
      else if (nm == "energy") r[0] = 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + u[3] * u[3] + u[4] * u[4] + u[5] * u[5]);
      else if (nm == "error") {
        double s2 = 0.0;
        for (int k = 0; k < 3; k++) s2 += (u[k] - exact[k]) * (u[k] - exact[k]);
        r[0] = std::sqrt(s2);
      }
    }
    return 0;
  };
  const int ns = dim == 2 ? 3 : 4;
  for (size_t vi = 0; vi < views.size(); vi++) {
    // <prefix>-sample<iteration><field>.pos (DGSensor.cc:512-523)
    char name[1024];
    std::snprintf(name, sizeof(name), "%s-sample%d%s.pos", prefix, iteration, views[vi].name);
    File out(name, "w");
    if (!out.f) return fail(c, DG_ERR_ARG, "cannot write %s", name);
    std::fprintf(out.f, "View \"Exported field\" {\n");
    const int nc = views[vi].ncomp;
    const char* tag = dim == 2 ? (nc == 3 ? "VT" : "ST") : (nc == 3 ? "VS" : "SS");
    for (int e = 0; e < E; e++)
      for (const auto& sb : sub) {
        std::fprintf(out.f, "%s (", tag);
        for (int k = 0; k < ns; k++) {
          const double* x = &xq[((size_t)e * NN + sb[k]) * 3];
          std::fprintf(out.f, "%s%.10g,%.10g,%.10g", k ? "," : "", x[0], x[1], x[2]);
        }
        std::fprintf(out.f, ") {");
        for (int k = 0; k < ns; k++) {
          double r[3] = {0, 0, 0};
          if (point_value((int)vi, e, sb[k], r)) return fail(c, DG_ERR_ARG, "unknown exact field %d", exact_field_id);
          for (int j = 0; j < nc; j++) std::fprintf(out.f, "%s%.10g", k == 0 && j == 0 ? " " : ",", r[j]);
        }
        std::fprintf(out.f, "};\n");
      }
    std::fprintf(out.f, "};\n");
    if (std::ferror(out.f)) return fail(c, DG_ERR_ARG, "write error on %s", name);
  }
  return DG_OK;
}

}  // extern "C"
