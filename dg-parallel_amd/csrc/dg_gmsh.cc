// Gmsh version-1 mesh files ($NOD/$NOE nodes, $ELM elements) into a mesh
// descriptor: the path's mesh ingest, replacing mImportExport::importGmshFile
// (mImportExport.cc:77-170) and mMesh::setPeriodicBC (mMesh.cc:457-523) for
// straight-sided triangles and tetrahedra.
//
// Element records are "id type physical elementary nnodes nodes...".
//  - 2D files: triangles (type 2) are the elements; lines (type 1) carry the
//    physical tag of the boundary edges they cover.  Quadrilaterals (type 3) are
//    rejected, or split along their (v0, v2) diagonal on request.
//  - 3D files: tetrahedra (type 4) are the elements; triangles (type 2) carry the
//    tags of boundary faces.
//  - Points (type 15) are ignored; any other type is rejected.
// Physical tag 100 marks interior entities and is ignored.  Boundary edges tagged
// 510 are paired by equal y coordinates of their ends, 610 by equal x (within
// 1e-10, the reference's tolerance); the vertices of each pair are given one
// topological id, which makes the two sides face neighbours (translation
// periodicity).  Every other tag goes to the bc array of the element face it
// covers; untagged faces without neighbour get DG_TAG_FARFIELD.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dg_capi.h"

namespace {

struct GmshMesh {
  int dim = 0;
  std::vector<int32_t> elems;         // [E][dim+1] node indices
  std::vector<double> nodes;          // [N][3]
  std::vector<long> ids;              // [N] node numbers of the file
  std::vector<std::vector<int>> bnd;  // tagged boundary entities: node indices, tag last
  int split = 0;
};

int find_root(std::vector<int>& p, int x) {
  while (p[x] != x) x = p[x] = p[p[x]];
  return x;
}

dg_status parse(const char* path, int split_quads, GmshMesh& M, std::string& err) {
  FILE* f = std::fopen(path, "r");
  if (!f) {
    err = std::string("cannot open ") + path;
    return DG_ERR_ARG;
  }
  std::unordered_map<long, int> node_at;  // node id -> index
  M.ids.clear();
  std::vector<std::vector<long>> tris, tets, lines, quads;
  std::vector<int> tri_tag, line_tag;
  char w[256];
  dg_status st = DG_OK;
  bool have_nodes = false, have_elems = false;
  while (st == DG_OK && fscanf(f, "%255s", w) == 1) {
    if (!std::strcmp(w, "$NOD") || !std::strcmp(w, "$NOE")) {
      long n = 0;
      if (fscanf(f, "%ld", &n) != 1 || n <= 0) {
        st = DG_ERR_ARG;
        err = "bad node count";
        break;
      }
      M.nodes.resize((size_t)n * 3);
      for (long i = 0; i < n; i++) {
        long id;
        double x, y, z;
        if (fscanf(f, "%ld %lf %lf %lf", &id, &x, &y, &z) != 4) {
          st = DG_ERR_ARG;
          err = "truncated node list";
          break;
        }
        node_at[id] = (int)i;
        M.ids.push_back(id);
        M.nodes[i * 3] = x;
        M.nodes[i * 3 + 1] = y;
        M.nodes[i * 3 + 2] = z;
      }
      have_nodes = true;
    } else if (!std::strcmp(w, "$ELM")) {
      long n = 0;
      if (fscanf(f, "%ld", &n) != 1 || n < 0) {
        st = DG_ERR_ARG;
        err = "bad element count";
        break;
      }
      for (long i = 0; i < n && st == DG_OK; i++) {
        long id, nn;
        int type, phys, elem;
        if (fscanf(f, "%ld %d %d %d %ld", &id, &type, &phys, &elem, &nn) != 5 || nn < 1 || nn > 64) {
          st = DG_ERR_ARG;
          err = "bad element record";
          break;
        }
        std::vector<long> v(nn);
        for (long k = 0; k < nn; k++)
          if (fscanf(f, "%ld", &v[k]) != 1) {
            st = DG_ERR_ARG;
            err = "truncated element record";
          }
        if (st != DG_OK) break;
        if (type == 15) continue;
        if (type == 1 && nn == 2) {
          lines.push_back(v);
          line_tag.push_back(phys);
        } else if (type == 2 && nn == 3) {
          tris.push_back(v);
          tri_tag.push_back(phys);
        } else if (type == 3 && nn == 4) {
          quads.push_back(v);
        } else if (type == 4 && nn == 4) {
          tets.push_back(v);
        } else {
          st = DG_ERR_UNSUPPORTED;
          char b[128];
          std::snprintf(b, sizeof(b), "element type %d with %ld nodes is not built (straight triangles and tetrahedra)",
                        type, nn);
          err = b;
        }
      }
      have_elems = true;
    }
  }
  std::fclose(f);
  if (st != DG_OK) return st;
  if (!have_nodes || !have_elems) {
    err = "no $NOD/$NOE or $ELM section";
    return DG_ERR_ARG;
  }
  auto idx = [&](long id, int& out) {
    auto it = node_at.find(id);
    if (it == node_at.end()) return false;
    out = it->second;
    return true;
  };
  M.dim = tets.empty() ? 2 : 3;
  if (M.dim == 3) {
    if (!quads.empty()) {
      err = "quadrilaterals in a tetrahedral mesh";
      return DG_ERR_UNSUPPORTED;
    }
    for (auto& t : tets)
      for (long id : t) {
        int k;
        if (!idx(id, k)) {
          err = "element refers to a missing node";
          return DG_ERR_ARG;
        }
        M.elems.push_back(k);
      }
    for (size_t i = 0; i < tris.size(); i++) {
      std::vector<int> b;
      for (long id : tris[i]) {
        int k;
        if (!idx(id, k)) {
          err = "boundary face refers to a missing node";
          return DG_ERR_ARG;
        }
        b.push_back(k);
      }
      b.push_back(tri_tag[i]);
      M.bnd.push_back(b);
    }
    return DG_OK;
  }
  if (!quads.empty() && !split_quads) {
    err = "quadrilaterals are not built (pass split_quads to split each along its (v0, v2) diagonal)";
    return DG_ERR_UNSUPPORTED;
  }
  auto push_tri = [&](long a, long b, long c) {
    for (long id : {a, b, c}) {
      int k;
      if (!idx(id, k)) return false;
      M.elems.push_back(k);
    }
    return true;
  };
  for (auto& t : tris)
    if (!push_tri(t[0], t[1], t[2])) {
      err = "element refers to a missing node";
      return DG_ERR_ARG;
    }
  for (auto& q : quads) {
    if (!push_tri(q[0], q[1], q[2]) || !push_tri(q[0], q[2], q[3])) {
      err = "element refers to a missing node";
      return DG_ERR_ARG;
    }
    M.split++;
  }
  for (size_t i = 0; i < lines.size(); i++) {
    int a, b;
    if (!idx(lines[i][0], a) || !idx(lines[i][1], b)) {
      err = "boundary edge refers to a missing node";
      return DG_ERR_ARG;
    }
    M.bnd.push_back({a, b, line_tag[i]});
  }
  return DG_OK;
}

}  // namespace

extern "C" dg_status dg_mesh_read_gmsh(const char* path, int split_quads, dg_gmsh_info* info, int32_t* vids,
                                       double* xyz, int32_t* bc, char* err_out, size_t err_len) {
  std::string err;
  auto fail = [&](dg_status s) {
    if (err_out && err_len) std::snprintf(err_out, err_len, "%s", err.c_str());
    return s;
  };
  if (!path || !info) {
    err = "null argument";
    return fail(DG_ERR_ARG);
  }
  GmshMesh M;
  dg_status st = parse(path, split_quads, M, err);
  if (st != DG_OK) return fail(st);
  const int nv = M.dim + 1, N = (int)(M.nodes.size() / 3);
  const int E = (int)(M.elems.size() / nv);
  if (E <= 0) {
    err = "no elements";
    return fail(DG_ERR_ARG);
  }
  // periodic pairs (2D): 510 by equal y, 610 by equal x; one id per vertex class
  std::vector<int> root(N);
  std::iota(root.begin(), root.end(), 0);
  int pairs = 0;
  const double SN = 1e-10;
  if (M.dim == 2) {
    for (int tag : {510, 610}) {
      const int c = tag == 510 ? 1 : 0;  // the coordinate that must match
      std::vector<std::pair<std::pair<double, double>, int>> es;
      for (size_t i = 0; i < M.bnd.size(); i++)
        if (M.bnd[i][2] == tag) {
          const double p = M.nodes[M.bnd[i][0] * 3 + c], q = M.nodes[M.bnd[i][1] * 3 + c];
          es.push_back({{std::min(p, q), std::max(p, q)}, (int)i});
        }
      // sorted by their (min, max) coordinate, partners are neighbours
      std::sort(es.begin(), es.end());
      for (size_t i = 0; i < es.size(); i += 2) {
        const size_t j = i + 1;
        if (j >= es.size() || std::fabs(es[j].first.first - es[i].first.first) >= SN ||
            std::fabs(es[j].first.second - es[i].first.second) >= SN) {
          char b[160];
          std::snprintf(b, sizeof(b), "periodic edge tagged %d at %c in [%g, %g] has no partner", tag, c ? 'y' : 'x',
                        es[i].first.first, es[i].first.second);
          err = b;
          return fail(DG_ERR_ARG);
        }
        const auto& A = M.bnd[es[i].second];
        const auto& B = M.bnd[es[j].second];
        const bool direct = std::fabs(M.nodes[A[0] * 3 + c] - M.nodes[B[0] * 3 + c]) < SN;
        const int a0 = A[0], a1 = A[1], b0 = direct ? B[0] : B[1], b1 = direct ? B[1] : B[0];
        root[find_root(root, a0)] = find_root(root, b0);
        root[find_root(root, a1)] = find_root(root, b1);
        pairs++;
      }
    }
  }
  // tags of the boundary entities by their sorted node classes
  std::map<std::vector<int>, int> tag_of;
  for (const auto& b : M.bnd) {
    const int tag = b.back();
    if (tag == 100 || (M.dim == 2 && (tag == 510 || tag == 610))) continue;
    std::vector<int> key(b.begin(), b.end() - 1);
    for (int& k : key) k = find_root(root, k);
    std::sort(key.begin(), key.end());
    tag_of[key] = tag;
  }
  // the vertex id of a class: its smallest node number, minus one (node numbers
  // of a file written by dg_mesh_write_gmsh are vertex id + 1)
  std::vector<long> cls(N, LONG_MAX);
  for (int i = 0; i < N; i++) {
    const int r = find_root(root, i);
    cls[r] = std::min(cls[r], M.ids[i]);
  }
  for (int i = 0; i < N; i++)
    if (root[i] == i && (cls[i] - 1 < 0 || cls[i] - 1 > INT32_MAX - 1)) {
      err = "node numbers must lie in 1..2^31-1";
      return fail(DG_ERR_ARG);
    }
  info->dim = M.dim;
  info->n_elems = E;
  info->n_nodes = N;
  info->n_quads_split = M.split;
  info->n_periodic_pairs = pairs;
  info->n_tagged = (int)tag_of.size();
  if (!vids && !xyz && !bc) return DG_OK;  // sizes only
  // local face l of a triangle is edge (v_l, v_{l+1}); of a tetrahedron the face
  // without vertex (l + 3) % 4, i.e. {0,1,2}, {0,1,3}, {1,2,3}, {0,2,3}
  static const int TETF[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};
  for (int e = 0; e < E; e++) {
    for (int v = 0; v < nv; v++) {
      const int k = M.elems[(size_t)e * nv + v];
      if (vids) vids[(size_t)e * nv + v] = (int32_t)(cls[find_root(root, k)] - 1);
      if (xyz)
        for (int d = 0; d < M.dim; d++) xyz[((size_t)e * nv + v) * M.dim + d] = M.nodes[(size_t)k * 3 + d];
    }
    if (bc)
      for (int l = 0; l < nv; l++) {
        std::vector<int> key;
        if (M.dim == 2) {
          key = {M.elems[(size_t)e * 3 + l], M.elems[(size_t)e * 3 + (l + 1) % 3]};
        } else {
          for (int i = 0; i < 3; i++) key.push_back(M.elems[(size_t)e * 4 + TETF[l][i]]);
        }
        for (int& k : key) k = find_root(root, k);
        std::sort(key.begin(), key.end());
        auto it = tag_of.find(key);
        bc[(size_t)e * nv + l] = it == tag_of.end() ? DG_TAG_FARFIELD : it->second;
      }
  }
  return DG_OK;
}
