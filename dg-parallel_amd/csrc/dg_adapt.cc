// h-adaptivity on triangles: DGAnalysis::adaptH's mesh operations (DGPref.cc:
// 78-192) with mDGMesh::split / unsplit (mDGMesh.cc:234-352, 463-552) and the
// state transfer of DGAnalysis::splitCell / unsplitCell (DGPref.cc:632-671).
// The caller marks the cells (the reference picks them from its error
// estimate, DGPref.cc:27-76); this file refines and coarsens the caller's mesh
// on the host, keeps it 1-irregular, transfers the state by exact L2
// projection, and builds the context anew on the new mesh (faces of half an
// edge run the hanging-face pass, dg_hang.h).
#include <algorithm>
#include <cstring>
#include <vector>

#include "dg_ctx.h"
#include "dg_tables.h"

namespace {

// children of a triangle (P0, P1, P2) with edge midpoints M01, M12, M20, as
// corners in the parent's reference coordinates: (P0, M01, M20),
// (M01, P1, M12), (M20, M12, P2), (M12, M20, M01)
const double KID[4][3][2] = {{{0, 0}, {0.5, 0}, {0, 0.5}},
                             {{0.5, 0}, {1, 0}, {0.5, 0.5}},
                             {{0, 0.5}, {0.5, 0.5}, {0, 1}},
                             {{0.5, 0.5}, {0, 0.5}, {0.5, 0}}};

// T[c][k][j] = int_ref phi_k(xi) phi_j(chi_c(xi)): the L2 projection onto child
// c of the parent's mode j (chi_c maps the child's reference triangle onto its
// part of the parent's); exact with the rule of degree 2P
std::vector<double> transfer(int P) {
  const int NP = np_tri(P);
  std::vector<double> T((size_t)4 * NP * NP, 0.0);
  const auto rule = tri_rule(2 * P);
  double phi[64], phj[64];
  for (int c = 0; c < 4; c++)
    for (const auto& q : rule) {
      basis(2, P, q.x, phi, nullptr);
      double xp[3] = {0, 0, 0};
      for (int d = 0; d < 2; d++)
        xp[d] = KID[c][0][d] + q.x[0] * (KID[c][1][d] - KID[c][0][d]) + q.x[1] * (KID[c][2][d] - KID[c][0][d]);
      basis(2, P, xp, phj, nullptr);
      for (int k = 0; k < NP; k++)
        for (int j = 0; j < NP; j++) T[((size_t)c * NP + k) * NP + j] += q.w * phi[k] * phj[j];
    }
  return T;
}

}  // namespace

extern "C" {

dg_status dg_adapt_h(dg_ctx* c, const int8_t* marks, int max_level, int32_t* n_elems_out) {
  if (!ready(c) || !marks || max_level < 0) return fail(c, DG_ERR_ARG, "dg_adapt_h: bad argument");
  if (c->dim != 2 || c->mixed) return fail(c, DG_ERR_UNSUPPORTED, "h-adaptivity runs triangles of one order");
  if (c->prec != DG_PREC_FP64 || c->kernel != 0)
    return fail(c, DG_ERR_UNSUPPORTED, "h-adaptivity runs the FP64 form of the tiled kernel");
  if (c->pvel) return fail(c, DG_ERR_UNSUPPORTED, "h-adaptivity runs meshes without a velocity field");
  const int E = c->E;
  if (c->h_level.empty()) {
    c->h_level.assign(E, 0);
    c->h_fam.assign(E, -1);
    c->h_ci.assign(E, 0);
  }
  if (c->h_nextid < 0) {
    int32_t mx = 0;
    for (int32_t v : c->m_vids) mx = std::max(mx, v);
    c->h_nextid = mx + 1;
  }
  const std::vector<int32_t>& lev = c->h_level;
  // face neighbours of the current mesh (halves of a split face included)
  std::vector<std::vector<int32_t>> nb(E);
  for (size_t i = 0; i < c->faces.size(); i += 4)
    if (c->faces[i + 2] >= 0) {
      nb[c->faces[i]].push_back(c->faces[i + 2]);
      nb[c->faces[i + 2]].push_back(c->faces[i]);
    }
  // splits: the marked cells below max_level, closed so that no cell is split
  // beside a coarser neighbour (its halves would be quartered)
  std::vector<char> split(E, 0);
  std::vector<int32_t> work;
  for (int e = 0; e < E; e++)
    if (marks[e] > 0 && lev[e] < max_level) {
      split[e] = 1;
      work.push_back(e);
    }
  while (!work.empty()) {
    const int e = work.back();
    work.pop_back();
    for (int n : nb[e])
      if (!split[n] && lev[n] < lev[e]) {
        split[n] = 1;
        work.push_back(n);
      }
  }
  // merges: families whose four children are all current, marked < 0 and not
  // split, and whose outside neighbours stay within one level of the parent
  const int nfam = (int)c->hfam.size();
  std::vector<std::vector<int32_t>> kids(nfam);  // current children of each family
  for (int e = 0; e < E; e++)
    if (c->h_fam[e] >= 0) kids[c->h_fam[e]].push_back(e);
  std::vector<char> merge(nfam, 0);
  for (int F = 0; F < nfam; F++) {
    if (kids[F].size() != 4) continue;
    bool ok = true;
    for (int e2 : kids[F]) ok = ok && marks[e2] < 0 && !split[e2];
    if (!ok) continue;
    const int L = lev[kids[F][0]];  // the children's level; the parent's is L - 1
    for (int e2 : kids[F])
      for (int n : nb[e2])
        if (c->h_fam[n] != F && lev[n] + (split[n] ? 1 : 0) > L) ok = false;
    if (ok) merge[F] = 1;
  }
  // the new mesh in order: a split cell's four children in its place, a merged
  // family's parent in place of its first child
  const int P = c->P, NP = np_tri(P), NF = c->Nf, BS = NF * NP;
  std::vector<double> U((size_t)E * BS);
  dg_status st = dg_get_state(c, U.data(), U.size());
  if (st) return st;
  const std::vector<double> T = transfer(P);
  std::vector<int32_t> nv, nbc, nlev, nfam2, nci;
  std::vector<double> nx, nU;
  auto mid = [&](int32_t a, int32_t b) {
    const uint64_t k = (uint64_t)(uint32_t)std::min(a, b) << 32 | (uint32_t)std::max(a, b);
    auto it = c->h_mid.find(k);
    if (it != c->h_mid.end()) return it->second;
    const int32_t id = c->h_nextid++;
    c->h_mid[k] = id;
    return id;
  };
  for (int e = 0; e < E; e++) {
    const int32_t* V = c->m_vids.data() + (size_t)e * 3;
    const double* X = c->xyz.data() + (size_t)e * 6;
    const int32_t tag[3] = {c->m_bc.empty() ? 0 : c->m_bc[e * 3], c->m_bc.empty() ? 0 : c->m_bc[e * 3 + 1],
                            c->m_bc.empty() ? 0 : c->m_bc[e * 3 + 2]};
    const double* Ue = U.data() + (size_t)e * BS;
    const int F = c->h_fam[e];
    if (F >= 0 && merge[F]) {
      if (e != kids[F][0]) continue;
      // the parent: 1/4 sum_c T_c^T U_c (children in any order, by place)
      const dg_ctx::HFam fam = c->hfam[F];
      std::vector<double> up(BS, 0.0);
      for (int e2 : kids[F]) {
        const int ci = c->h_ci[e2];
        const double* Uc = U.data() + (size_t)e2 * BS;
        for (int f = 0; f < NF; f++)
          for (int j = 0; j < NP; j++) {
            double sacc = 0.0;
            for (int k = 0; k < NP; k++) sacc += T[((size_t)ci * NP + k) * NP + j] * Uc[f * NP + k];
            up[f * NP + j] += 0.25 * sacc;
          }
      }
      nv.insert(nv.end(), fam.vids, fam.vids + 3);
      nx.insert(nx.end(), fam.xyz, fam.xyz + 6);
      nbc.insert(nbc.end(), fam.bc, fam.bc + 3);
      nlev.push_back(fam.level);
      nfam2.push_back(fam.fam);
      nci.push_back(fam.ci);
      nU.insert(nU.end(), up.begin(), up.end());
      continue;
    }
    if (!split[e]) {
      nv.insert(nv.end(), V, V + 3);
      nx.insert(nx.end(), X, X + 6);
      nbc.insert(nbc.end(), tag, tag + 3);
      nlev.push_back(lev[e]);
      nfam2.push_back(F);
      nci.push_back(c->h_ci[e]);
      nU.insert(nU.end(), Ue, Ue + BS);
      continue;
    }
    // red refinement: children (P0, M01, M20), (M01, P1, M12), (M20, M12, P2),
    // (M12, M20, M01); midpoint coordinates from this cell's own corners
    dg_ctx::HFam fam;
    std::memcpy(fam.vids, V, sizeof(fam.vids));
    std::memcpy(fam.xyz, X, sizeof(fam.xyz));
    std::memcpy(fam.bc, tag, sizeof(fam.bc));
    fam.level = lev[e];
    fam.fam = F;
    fam.ci = c->h_ci[e];
    const int newF = (int)c->hfam.size();
    c->hfam.push_back(fam);
    const int32_t m01 = mid(V[0], V[1]), m12 = mid(V[1], V[2]), m20 = mid(V[2], V[0]);
    const int32_t cv[4][3] = {{V[0], m01, m20}, {m01, V[1], m12}, {m20, m12, V[2]}, {m12, m20, m01}};
    const int32_t cb[4][3] = {{tag[0], 0, tag[2]}, {tag[0], tag[1], 0}, {0, tag[1], tag[2]}, {0, 0, 0}};
    for (int ci = 0; ci < 4; ci++) {
      nv.insert(nv.end(), cv[ci], cv[ci] + 3);
      for (int v = 0; v < 3; v++)
        for (int d = 0; d < 2; d++) {
          const double r0 = KID[ci][v][0], r1 = KID[ci][v][1];
          nx.push_back(X[d] + r0 * (X[2 + d] - X[d]) + r1 * (X[4 + d] - X[d]));
        }
      nbc.insert(nbc.end(), cb[ci], cb[ci] + 3);
      nlev.push_back(lev[e] + 1);
      nfam2.push_back(newF);
      nci.push_back(ci);
      for (int f = 0; f < NF; f++)
        for (int k = 0; k < NP; k++) {
          double sacc = 0.0;
          for (int j = 0; j < NP; j++) sacc += T[((size_t)ci * NP + k) * NP + j] * Ue[f * NP + j];
          nU.push_back(sacc);
        }
    }
  }
  // the merged families' children leave the mesh; their families stay listed
  std::vector<int32_t> old_level = c->h_level, old_fam = c->h_fam, old_ci = c->h_ci;
  c->h_level = nlev;
  c->h_fam = nfam2;
  c->h_ci = nci;
  const bool had_bc = !c->m_bc.empty();
  st = ctx_remesh(c, nv, nx, had_bc ? nbc : std::vector<int32_t>(), nU);
  if (st) {
    // the context keeps its mesh: its history too
    c->h_level.swap(old_level);
    c->h_fam.swap(old_fam);
    c->h_ci.swap(old_ci);
    return st;
  }
  if (n_elems_out) *n_elems_out = c->E;
  return DG_OK;
}

dg_status dg_get_mesh(const dg_ctx* c, int32_t* n_elems, int32_t* vids, double* xyz, int32_t* bc, int32_t* levels) {
  if (!c || !c->stream || c->dim != 2) return DG_ERR_ARG;
  if (n_elems) *n_elems = c->E;
  if (vids) std::memcpy(vids, c->m_vids.data(), (size_t)c->E * 3 * 4);
  if (xyz) std::memcpy(xyz, c->xyz.data(), (size_t)c->E * 6 * 8);
  if (bc)
    for (int i = 0; i < c->E * 3; i++) bc[i] = c->m_bc.empty() ? 0 : c->m_bc[i];
  if (levels)
    for (int e = 0; e < c->E; e++) levels[e] = c->h_level.empty() ? 0 : c->h_level[e];
  return DG_OK;
}

}  // extern "C"
