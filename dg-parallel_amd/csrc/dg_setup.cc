// Host-side FP64 table construction for the triangle kernels.
#include "dg_setup.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "dg_tables.h"

namespace dgp {

static dg_status failS(std::string& err, dg_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  err = buf;
  return s;
}

// faces of half an edge among the faces without a neighbour (see dg_setup.h).
// A face split into more than two pieces is not found (its pieces are no
// halves): such faces stay boundary faces; dg_adapt_h keeps meshes 1-irregular.
static void find_hanging(const dg_mesh_desc* m, TriTopo& T) {
  T.hang.clear();
  const int E = m->n_elems;
  // faces without a neighbour: (vertex, e*3+l) for both ends, sorted
  std::vector<std::pair<int32_t, int32_t>> at;
  for (int e = 0; e < E; e++)
    for (int l = 0; l < 3; l++)
      if (T.nbe[e * 3 + l] < 0) {
        at.push_back({m->vids[e * 3 + l], e * 3 + l});
        at.push_back({m->vids[e * 3 + (l + 1) % 3], e * 3 + l});
      }
  if (at.empty()) return;
  std::sort(at.begin(), at.end());
  auto range = [&](int32_t v) {
    auto lo = std::lower_bound(at.begin(), at.end(), std::make_pair(v, (int32_t)INT32_MIN));
    auto hi = std::upper_bound(at.begin(), at.end(), std::make_pair(v, (int32_t)INT32_MAX));
    return std::make_pair(lo, hi);
  };
  auto vert = [&](int f, int i) { return m->vids[(f / 3) * 3 + (f % 3 + i) % 3]; };
  // vector from the face's vertex i to vertex j (0 or 1) in its element's coordinates
  auto vec = [&](int f, int i, int j, double* d) {
    const double* X = m->xyz + (size_t)(f / 3) * 6;
    const int a = (f % 3 + i) % 3, b = (f % 3 + j) % 3;
    d[0] = X[b * 2] - X[a * 2];
    d[1] = X[b * 2 + 1] - X[a * 2 + 1];
  };
  std::vector<char> used((size_t)E * 3, 0);
  std::vector<std::pair<int, int>> halves;  // per coarse face found: the two half faces
  for (int e = 0; e < E; e++)
    for (int l = 0; l < 3; l++) {
      const int F = e * 3 + l;
      if (T.nbe[F] >= 0 || used[F]) continue;
      const int32_t a = vert(F, 0), b = vert(F, 1);
      int h1 = -1, h2 = -1;
      auto ra = range(a);
      for (auto it = ra.first; it != ra.second && h1 < 0; ++it) {
        const int g = it->second;
        if (g / 3 == e || used[g]) continue;
        const int ia = vert(g, 0) == a ? 0 : 1;
        const int32_t x = vert(g, 1 - ia);
        auto rx = range(x);
        for (auto jt = rx.first; jt != rx.second; ++jt) {
          const int g2 = jt->second;
          if (g2 == g || g2 / 3 == e || used[g2]) continue;
          const int ix = vert(g2, 0) == x ? 0 : 1;
          if (vert(g2, 1 - ix) != b) continue;
          // x is the midpoint of (a, b): x - a equals b - x
          double d1[2], d2[2];
          vec(g, ia, 1 - ia, d1);
          vec(g2, ix, 1 - ix, d2);
          if (std::fabs(d1[0] - d2[0]) + std::fabs(d1[1] - d2[1]) > 1e-9 * (std::fabs(d1[0]) + std::fabs(d1[1])))
            continue;
          h1 = g;
          h2 = g2;
          break;
        }
      }
      if (h1 < 0) continue;
      used[F] = used[h1] = used[h2] = 1;
      for (int t = 0; t < 2; t++) {
        const int g = t ? h2 : h1;
        HangFace H;
        H.child = g / 3;
        H.lc = g % 3;
        H.K = e;
        H.lK = l;
        // the half holding K's vertex l (= a) is half 0
        const int32_t ga = vert(g, 0), gb = vert(g, 1);
        H.h = (ga == a || gb == a) ? 0 : 1;
        // K runs a -> b; the child runs ga -> gb: along K's direction when it
        // starts at a (half 0) or at the midpoint (half 1)
        const bool along = H.h == 0 ? ga == a : gb == b;
        H.rev = along ? 0 : 1;
        T.hang.push_back(H);
        T.bcode[g] = BC_HANG;
      }
      T.bcode[F] = BC_HANG;
    }
  // faces list: each coarse face becomes its two halves (child left, K right)
  std::vector<int32_t> nf;
  nf.reserve(T.faces.size() + 4 * T.hang.size());
  for (size_t i = 0; i < T.faces.size(); i += 4) {
    const int Le = T.faces[i], Ll = T.faces[i + 1];
    if (T.faces[i + 2] < 0 && used[Le * 3 + Ll]) continue;  // a coarse face or a half: listed below
    nf.insert(nf.end(), T.faces.begin() + i, T.faces.begin() + i + 4);
  }
  for (const HangFace& H : T.hang) nf.insert(nf.end(), {H.child, H.lc, H.K, H.lK});
  T.faces.swap(nf);
}

dg_status build_tri_topo(const dg_mesh_desc* m, TriTopo& T, std::string& err) {
  const int E = m->n_elems;
  T.E = E;
  // ---- connectivity (mMesh.cc:321-378, translation-periodic by vertex ids): edges
  //      keyed by their sorted vertex ids, ties in (element, local edge) order
  std::vector<std::pair<uint64_t, int32_t>> keys((size_t)E * 3);
  for (int e = 0; e < E; e++)
    for (int l = 0; l < 3; l++) {
      const uint32_t a = (uint32_t)m->vids[e * 3 + l], b = (uint32_t)m->vids[e * 3 + (l + 1) % 3];
      if (a == b) return failS(err, DG_ERR_ARG, "degenerate edge in element %d", e);
      keys[(size_t)e * 3 + l] = {(uint64_t)std::min(a, b) << 32 | std::max(a, b), e * 3 + l};
    }
  std::sort(keys.begin(), keys.end());
  T.nbe.assign((size_t)E * 3, -1);
  T.nbl.assign((size_t)E * 3, -1);
  T.same.assign((size_t)E * 3, 0);
  T.faces.clear();
  T.faces.reserve((size_t)E * 3 * 4 / 2 + 16);
  for (size_t i = 0; i < keys.size();) {
    size_t j = i + 1;
    while (j < keys.size() && keys[j].first == keys[i].first) j++;
    if (j - i > 2)
      return failS(err, DG_ERR_ARG, "edge (%u,%u) shared by %zu elements", (unsigned)(keys[i].first >> 32),
                   (unsigned)(keys[i].first & 0xffffffffu), j - i);
    const int Le = keys[i].second / 3, Ll = keys[i].second % 3;
    if (j - i == 2) {
      const int Re = keys[i + 1].second / 3, Rl = keys[i + 1].second % 3;
      T.nbe[Le * 3 + Ll] = Re;
      T.nbl[Le * 3 + Ll] = Rl;
      T.nbe[Re * 3 + Rl] = Le;
      T.nbl[Re * 3 + Rl] = Ll;
      const int sm = m->vids[Le * 3 + Ll] == m->vids[Re * 3 + Rl];
      T.same[Le * 3 + Ll] = T.same[Re * 3 + Rl] = sm;
      T.faces.insert(T.faces.end(), {Le, Ll, Re, Rl});
    } else {
      T.faces.insert(T.faces.end(), {Le, Ll, -1, -1});
    }
    i = j;
  }
  T.bcode.assign((size_t)E * 3, 0);
  find_hanging(m, T);
  T.sig.assign(E, 0);
  for (int e = 0; e < E; e++) {
    int sg = 0;
    for (int l = 0; l < 3; l++) {
      int v;
      if (T.nbe[e * 3 + l] >= 0) {
        v = T.nbl[e * 3 + l] | (T.same[e * 3 + l] << 2);
      } else {
        if (T.bcode[e * 3 + l] == BC_HANG) {
          v = 8 + BC_HANG;
          sg = sg * 16 + v;
          continue;
        }
        const int tag = m->bc ? m->bc[e * 3 + l] : DG_TAG_FARFIELD;
        // 20000, the reference's curved wall, is the 40000 wall on straight faces
        const int code = (tag == DG_TAG_WALL || tag == DG_TAG_CURVED_WALL) ? BC_WALL
                         : tag == DG_TAG_FARFIELD2                          ? BC_FARFIELD2
                         : tag == DG_TAG_CHARACTERISTIC                     ? BC_CHAR
                                                                            : BC_FARFIELD;
        T.bcode[e * 3 + l] = (int8_t)code;
        v = 8 + code;
      }
      sg = sg * 16 + v;
    }
    T.sig[e] = sg;
  }

  // ---- geometry (Mapping.cc:442-566, 687-712; DGCell.cc:1209-1229) ---------------
  T.mg.assign((size_t)4 * E, 0.0);
  T.detj.assign(E, 0.0);
  T.size.assign(E, 0.0);
  for (int e = 0; e < E; e++) {
    const double* X = m->xyz + (size_t)e * 6;
    double J00 = X[2] - X[0], J01 = X[4] - X[0];  // dx/du, dx/dv
    double J10 = X[3] - X[1], J11 = X[5] - X[1];  // dy/du, dy/dv
    double det = J00 * J11 - J01 * J10;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return failS(err, DG_ERR_ARG, "degenerate element %d", e);
    double ad = std::fabs(det);
    // |J| * (J^-1)[r][d]
    double inv[2][2] = {{J11 / det, -J01 / det}, {-J10 / det, J00 / det}};
    for (int r = 0; r < 2; r++)
      for (int d = 0; d < 2; d++) T.mg[(size_t)e * 4 + r * 2 + d] = ad * inv[r][d];
    T.detj[e] = ad;
    T.size[e] = tri_cell_size(X);
  }
  return DG_OK;
}

double tri_cell_size(const double* X) {
  const double a = std::sqrt((X[2] - X[0]) * (X[2] - X[0]) + (X[3] - X[1]) * (X[3] - X[1]));
  const double b = std::sqrt((X[4] - X[2]) * (X[4] - X[2]) + (X[5] - X[3]) * (X[5] - X[3]));
  const double cc = std::sqrt((X[0] - X[4]) * (X[0] - X[4]) + (X[1] - X[5]) * (X[1] - X[5]));
  const double hp = 0.5 * (a + b + cc);
  return std::sqrt(hp * (hp - a) * (hp - b) * (hp - cc)) / hp * 2.;
}

// f(i0, i1) over [0, n) split among up to 8 host threads (one thread below
// 16384 items); every range is handled by exactly one call.  A thread that
// cannot be started leaves its range, and the ones after it, to the caller's
// thread.
template <class F>
static void par_for(int n, F f) {
  const unsigned hc = std::thread::hardware_concurrency();
  const int nt = n < 16384 ? 1 : (int)std::min(8u, hc ? hc : 1u);
  if (nt <= 1) {
    f(0, n);
    return;
  }
  const int chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  int t = 1;
  for (; t < nt; t++) {
    const int i0 = t * chunk, i1 = std::min(n, i0 + chunk);
    if (i0 >= i1) break;
    try {
      th.emplace_back(f, i0, i1);
    } catch (...) {
      break;
    }
  }
  f(0, std::min(n, chunk));
  for (; t < nt; t++) {
    const int i0 = t * chunk, i1 = std::min(n, i0 + chunk);
    if (i0 < i1) f(i0, i1);
  }
  for (auto& x : th) x.join();
}

void tri_layout(const TriTopo& T, const int* ord, TriSetup& S, const int32_t* grp, bool geom) {
  const int E = T.E, NF = S.NF;
  S.mixed = false;
  for (int e = 1; e < E && !S.mixed; e++) S.mixed = ord[e] != ord[0];
  S.P = S.mixed ? 0 : ord[0];
  S.NP = S.mixed ? 0 : np_tri(ord[0]);
  S.coff.resize(E + 1);
  S.coff[0] = 0;
  for (int e = 0; e < E; e++) S.coff[e + 1] = S.coff[e] + (long)NF * np_tri(ord[e]);
  // order buckets, blocks of BLK caller elements, signatures: each block's
  // elements of one order are few, so a stable sort per block is cheap.  The
  // blocks' places in each bucket come from per-block counts, so the blocks
  // are placed independently (in parallel).
  const int BLK = 2 * 64;
  const int nblk = (E + BLK - 1) / BLK;
  std::vector<std::array<int, MAXP_TRI + 1>> bcnt(nblk);
  par_for(nblk, [&](int k0, int k1) {
    for (int k = k0; k < k1; k++) {
      bcnt[k].fill(0);
      for (int e = k * BLK; e < std::min(E, (k + 1) * BLK); e++) bcnt[k][ord[e]]++;
    }
  });
  int cnt[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < nblk; k++)
    for (int p = 1; p <= MAXP_TRI; p++) cnt[p] += bcnt[k][p];
  {
    // bcnt becomes each block's first slot in each order's run
    int at[6] = {0, 0, 0, 0, 0, 0};
    for (int p = 2; p <= MAXP_TRI; p++) at[p] = at[p - 1] + cnt[p - 1];
    for (int k = 0; k < nblk; k++)
      for (int p = 1; p <= MAXP_TRI; p++) {
        const int c = bcnt[k][p];
        bcnt[k][p] = at[p];
        at[p] += c;
      }
  }
  S.perm.resize(E);
  par_for(nblk, [&](int k0, int k1) {
    int el[MAXP_TRI + 1][BLK], sgs[MAXP_TRI + 1][BLK];
    for (int k = k0; k < k1; k++) {
      int n[MAXP_TRI + 1] = {0};
      for (int e = k * BLK; e < std::min(E, (k + 1) * BLK); e++) {
        // stable insertion by signature into this block's elements of order ord[e]
        const int p = ord[e], sg = T.sig[e];
        int i = n[p]++;
        for (; i > 0 && sgs[p][i - 1] > sg; i--) {
          sgs[p][i] = sgs[p][i - 1];
          el[p][i] = el[p][i - 1];
        }
        sgs[p][i] = sg;
        el[p][i] = e;
      }
      for (int p = 1; p <= MAXP_TRI; p++) std::copy(el[p], el[p] + n[p], S.perm.begin() + bcnt[k][p]);
    }
  });
  if (grp && !S.mixed) {
    // one order: groups first, then the same blocks and signatures
    for (int e = 0; e < E; e++) S.perm[e] = e;
    std::stable_sort(S.perm.begin(), S.perm.end(), [&](int a, int b) {
      if (grp[a] != grp[b]) return grp[a] < grp[b];
      if (a / BLK != b / BLK) return a / BLK < b / BLK;
      return T.sig[a] < T.sig[b];
    });
  }
  std::vector<int32_t> iperm(E);
  S.order.resize(E);
  par_for(E, [&](int i0, int i1) {
    for (int i = i0; i < i1; i++) {
      iperm[S.perm[i]] = i;
      S.order[i] = (int8_t)ord[S.perm[i]];
    }
  });
  S.bstart[0] = 0;
  S.boff[0] = 0;
  S.bstart[1] = 0;
  S.boff[1] = 0;
  for (int p = 1; p <= MAXP_TRI; p++) {
    S.bstart[p + 1] = S.bstart[p] + cnt[p];
    S.boff[p + 1] = S.boff[p] + (long)NF * np_tri(p) * cnt[p];
  }
  if (geom) {
    S.mg.resize((size_t)4 * E);
    S.detj.resize(E);
    S.size.resize(E);
  }
  S.link.resize((size_t)3 * E);
  par_for(E, [&](int i0, int i1) {
    for (int i = i0; i < i1; i++) {
      const int e = S.perm[i];
      if (geom) {
        for (int r = 0; r < 4; r++) S.mg[(size_t)r * E + i] = T.mg[(size_t)e * 4 + r];
        S.detj[i] = T.detj[e];
        S.size[i] = T.size[e];
      }
      for (int l = 0; l < 3; l++) {
        const int nb = T.nbe[e * 3 + l];
        S.link[(size_t)l * E + i] = nb >= 0 ? (iperm[nb] << 3) | (T.same[e * 3 + l] << 2) | T.nbl[e * 3 + l]
                                             : -(int32_t)T.bcode[e * 3 + l];
      }
    }
  });
}

void build_hang_tables(const TriSetup& S, std::vector<int32_t>& elem, std::vector<int32_t>& rstart,
                       std::vector<int32_t>& rec, std::vector<double>& ph) {
  const int E = S.E;
  std::vector<int32_t> iperm(E);
  for (int i = 0; i < E; i++) iperm[S.perm[i]] = i;
  std::vector<std::array<int32_t, 5>> r;  // own, local face, other, its local face, flags
  for (const HangFace& H : S.topo.hang) {
    const int fl = (H.rev ? 2 : 0) | (H.h << 2);
    r.push_back({iperm[H.child], H.lc, iperm[H.K], H.lK, fl});
    r.push_back({iperm[H.K], H.lK, iperm[H.child], H.lc, 1 | fl});
  }
  std::sort(r.begin(), r.end());
  for (size_t k = 0; k < r.size(); k++) {
    if (k == 0 || r[k][0] != r[k - 1][0]) {
      elem.push_back(r[k][0]);
      rstart.push_back((int32_t)k);
    }
    rec.insert(rec.end(), {r[k][1], r[k][2], r[k][3], r[k][4]});
  }
  rstart.push_back((int32_t)r.size());
  const int P = S.P, NP = np_tri(P), NQF = nqf_tri(P);
  ph.assign((size_t)3 * 2 * NQF * NP, 0.0);
  const auto gl = gl_rule(NQF);
  double phi[64], xi[3];
  for (int l = 0; l < 3; l++)
    for (int h = 0; h < 2; h++)
      for (int q = 0; q < NQF; q++) {
        // the face rule's point q lies at fraction (1 - s_q) / 2 from vertex l
        // (edge_point); on half h at h / 2 + (1 - s_q) / 4
        const double t = 0.5 * h + 0.25 * (1.0 - gl[q].x[0]);
        edge_point(l, 1.0 - 2.0 * t, xi);
        basis(2, P, xi, phi, nullptr);
        for (int k = 0; k < NP; k++) ph[((l * 2 + h) * NQF + q) * NP + k] = phi[k];
      }
}

dg_status build_tri_setup(const dg_mesh_desc* m, const dg_law_desc* law, const dg_options* opt, TriSetup& S,
                          std::string& err, const int32_t* grp) {
  if (!m || !law || !opt || !m->vids || !m->xyz || m->n_elems <= 0) return failS(err, DG_ERR_ARG, "null mesh/law/options");
  if (m->dim != 2) return failS(err, DG_ERR_UNSUPPORTED, "only triangles (dim 2) are built");
  S.E = m->n_elems;
  S.law = law->law;
  S.NF = law->law == DG_LAW_SCALAR      ? 1
         : law->law == DG_LAW_EULER2D   ? 4
         : law->law == DG_LAW_MAXWELL2D ? 3
         : law->law == DG_LAW_BURGERS2D ? 1
                                        : -1;
  if (S.NF < 0) return failS(err, DG_ERR_UNSUPPORTED, "law %d not built for triangles", law->law);
  if (S.law == DG_LAW_EULER2D && law->flux != DG_FLUX_LF && law->flux != DG_FLUX_ROE && law->flux != DG_FLUX_ROE_EF &&
      law->flux != DG_FLUX_EXACT)
    return failS(err, DG_ERR_UNSUPPORTED, "Euler flux %d not built", law->flux);
  if (S.law == DG_LAW_SCALAR && law->flux != DG_FLUX_UPWIND)
    return failS(err, DG_ERR_UNSUPPORTED, "scalar flux %d not built", law->flux);
  if (S.law == DG_LAW_MAXWELL2D && law->flux != DG_FLUX_UPWIND)
    return failS(err, DG_ERR_UNSUPPORTED, "Maxwell2d flux %d not built (its own upwind state: DG_FLUX_UPWIND)", law->flux);
  if (S.law == DG_LAW_BURGERS2D && law->flux != DG_FLUX_CENTRAL)
    return failS(err, DG_ERR_UNSUPPORTED, "Burgers2D flux %d not built (its central flux: DG_FLUX_CENTRAL)", law->flux);
  // per-element orders (mesh->order) or one order for all (opt->order)
  std::vector<int> ord(S.E, opt->order);
  if (m->order)
    for (int e = 0; e < S.E; e++) ord[e] = m->order[e];
  for (int e = 0; e < S.E; e++)
    if (ord[e] < 1 || ord[e] > MAXP_TRI)
      return failS(err, DG_ERR_UNSUPPORTED, "order %d of element %d outside 1..%d", ord[e], e, MAXP_TRI);
  bool mixed = false;
  for (int e = 1; e < S.E; e++)
    if (ord[e] != ord[0]) mixed = true;
  if (opt->precision != DG_PREC_FP64 && opt->precision != DG_PREC_FP32) return failS(err, DG_ERR_ARG, "bad precision");
  if (mixed && opt->precision != DG_PREC_FP64)
    return failS(err, DG_ERR_UNSUPPORTED, "mixed orders run in the FP64 form only");
  S.lp.gamma = law->gamma;
  for (int i = 0; i < 3; i++) S.lp.vel[i] = law->vel[i];
  for (int i = 0; i < 8; i++) S.lp.farfield[i] = law->farfield[i];
  S.lp.flux = law->flux;
  S.xyz.assign(m->xyz, m->xyz + (size_t)S.E * 6);
  dg_status st = build_tri_topo(m, S.topo, err);
  if (st != DG_OK) return st;
  // the characteristic far field (tag 30000) is an Euler boundary (EulerLaw.cc:
  // 185-278); the other laws' boundary fluxes give 30000
  // no meaning of its own and treat it as any far-field tag (ScalarLaw.cc:39-59,
  // MaxwellLaw.cc:115-130, Burgers2D.cc:31-49)
  S.lp.char_bc = 0;
  for (int8_t& b : S.topo.bcode)
    if (b == BC_CHAR) {
      if (S.law == DG_LAW_EULER2D)
        S.lp.char_bc = 1;
      else
        b = BC_FARFIELD;
    }
  S.faces = S.topo.faces;
  tri_layout(S.topo, ord.data(), S, grp);
  const int P = S.P;

  // ---- tables ---------------------------------------------------------------------
  {
    // every order's block (mixed-order kernels, limiter, sensor, dt)
    int o = 0;
    for (int p = 1; p <= MAXP_TRI; p++) {
      S.moff[p] = o;
      o += TriBlock::size(p);
    }
    S.mtab.assign(o, 0.0);
    for (int p = 1; p <= MAXP_TRI; p++) {
      try {
        build_tri_block(p, S.mtab.data() + S.moff[p]);
      } catch (const std::exception& ex) {
        return failS(err, DG_ERR_ARG, "%s", ex.what());
      }
      for (int nq = 0; nq < 6; nq++)
        S.ftoff[p][nq] = nq < p + 1 ? -1 : S.moff[p] + (nq == p + 1 ? TriBlock::off_phif(p) : TriBlock::off_phifx(p, nq));
    }
    for (int nq = 0; nq < 6; nq++) S.ftoff[0][nq] = -1;
    S.ntab.assign(MIX_NTAB, 0.0);
    for (int p = 1; p <= MAXP_TRI; p++)
      for (int nq = p + 1; nq <= MAXP_TRI + 1; nq++)
        std::copy_n(S.mtab.begin() + S.ftoff[p][nq], 3 * nq * np_tri(p), S.ntab.begin() + ntab_off(p, nq));
    S.glw.assign(GLW_OFF(MAXP_TRI + 2), 0.0);
    build_glw(S.glw.data());
    S.mvtx.assign((size_t)(MAXP_TRI + 1) * 4 * 15, 0.0);
    double z[3] = {0, 0, 0}, phi[64];
    static const double V[3][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
    for (int p = 1; p <= MAXP_TRI; p++) {
      basis(2, 0, z, phi, nullptr);
      S.mvtx[(size_t)p * 60] = phi[0];
      for (int v = 0; v < 3; v++) {
        basis(2, p, V[v], phi, nullptr);
        for (int k = 0; k < np_tri(p); k++) S.mvtx[(size_t)p * 60 + (1 + v) * 15 + k] = phi[k];
      }
    }
  }
  if (S.mixed) return DG_OK;
  const int NP = S.NP;
  S.tab.assign(TriBlock::size(P), 0.0);
  try {
    build_tri_block(P, S.tab.data());
  } catch (const std::exception& ex) {
    return failS(err, DG_ERR_ARG, "%s", ex.what());
  }
  S.tabf.assign(S.tab.begin(), S.tab.end());
  S.vtx.assign((size_t)4 * NP, 0.0);
  std::vector<double>& h_vtx = S.vtx;
  {
    double z[3] = {0, 0, 0}, phi[64];
    basis(2, 0, z, phi, nullptr);
    h_vtx[0] = phi[0];  // computeMean: coefficient 0 times phi_0 (DGCell.cc:1133-1137)
    static const double V[3][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
    for (int v = 0; v < 3; v++) {
      basis(2, P, V[v], phi, nullptr);
      for (int k = 0; k < NP; k++) h_vtx[(1 + v) * NP + k] = phi[k];
    }
  }

  return DG_OK;
}

// ---------------------------------------------------------------------------
// tetrahedra (mMesh.cc:321-378 with faces keyed by their sorted vertex triples;
// Mapping.cc:442-566 for the affine Jacobian)
dg_status build_tet_setup(const dg_mesh_desc* m, const dg_law_desc* law, const dg_options* opt, TetSetup& S,
                          std::string& err) {
  if (!m || !law || !opt || !m->vids || !m->xyz || m->n_elems <= 0) return failS(err, DG_ERR_ARG, "null mesh/law/options");
  if (law->law != DG_LAW_MAXWELL3D) return failS(err, DG_ERR_UNSUPPORTED, "tetrahedra run the Maxwell law only");
  if (law->flux != DG_FLUX_UPWIND && law->flux != DG_FLUX_CENTRAL)
    return failS(err, DG_ERR_UNSUPPORTED, "Maxwell flux %d not built", law->flux);
  if (opt->precision != DG_PREC_FP64 && opt->precision != DG_PREC_FP32) return failS(err, DG_ERR_ARG, "bad precision");
  int P = opt->order;
  if (m->order)
    for (int e = 0; e < m->n_elems; e++)
      if (m->order[e] != m->order[0]) return failS(err, DG_ERR_UNSUPPORTED, "mixed orders are built for triangles only");
  if (m->order) P = m->order[0];
  if (P < 1 || P > MAXP_TET) return failS(err, DG_ERR_UNSUPPORTED, "tet order %d outside 1..%d", P, MAXP_TET);
  S.E = m->n_elems;
  S.P = P;
  S.NP = np_tet(P);
  S.NF = 6;
  S.law = law->law;
  S.lp.gamma = law->gamma;
  for (int i = 0; i < 3; i++) S.lp.vel[i] = law->vel[i];
  for (int i = 0; i < 8; i++) S.lp.farfield[i] = law->farfield[i];
  S.lp.flux = law->flux;
  const int E = S.E;
  S.xyz.assign(m->xyz, m->xyz + (size_t)E * 12);

  struct FK {
    int32_t k[3];
    int32_t e, l;
    bool operator<(const FK& o) const {
      for (int i = 0; i < 3; i++)
        if (k[i] != o.k[i]) return k[i] < o.k[i];
      if (e != o.e) return e < o.e;
      return l < o.l;
    }
  };
  std::vector<FK> keys((size_t)E * 4);
  for (int e = 0; e < E; e++)
    for (int l = 0; l < 4; l++) {
      FK& k = keys[(size_t)e * 4 + l];
      for (int i = 0; i < 3; i++) k.k[i] = m->vids[e * 4 + TET_FACE_V[l][i]];
      std::sort(k.k, k.k + 3);
      if (k.k[0] == k.k[1] || k.k[1] == k.k[2]) return failS(err, DG_ERR_ARG, "degenerate face in element %d", e);
      k.e = e;
      k.l = l;
    }
  std::sort(keys.begin(), keys.end());
  std::vector<int32_t> nbe((size_t)E * 4, -1), nbl((size_t)E * 4, -1), nbo((size_t)E * 4, 0);
  auto orient = [&](int e, int l, int n, int ln) {
    // permutation o with vid_n[f(ln, PERM3[o][i])] == vid_e[f(l, i)]
    for (int o = 0; o < 6; o++) {
      bool ok = true;
      for (int i = 0; i < 3; i++)
        if (m->vids[n * 4 + TET_FACE_V[ln][PERM3[o][i]]] != m->vids[e * 4 + TET_FACE_V[l][i]]) ok = false;
      if (ok) return o;
    }
    return -1;
  };
  for (size_t i = 0; i < keys.size();) {
    size_t j = i + 1;
    while (j < keys.size() && keys[j].k[0] == keys[i].k[0] && keys[j].k[1] == keys[i].k[1] && keys[j].k[2] == keys[i].k[2]) j++;
    if (j - i > 2) return failS(err, DG_ERR_ARG, "face shared by %zu elements", j - i);
    const FK& L = keys[i];
    if (j - i == 2) {
      const FK& R = keys[i + 1];
      nbe[L.e * 4 + L.l] = R.e;
      nbl[L.e * 4 + L.l] = R.l;
      nbe[R.e * 4 + R.l] = L.e;
      nbl[R.e * 4 + R.l] = L.l;
      int o1 = orient(L.e, L.l, R.e, R.l), o2 = orient(R.e, R.l, L.e, L.l);
      if (o1 < 0 || o2 < 0) return failS(err, DG_ERR_ARG, "unmatched face vertices");
      nbo[L.e * 4 + L.l] = o1;
      nbo[R.e * 4 + R.l] = o2;
      S.faces.insert(S.faces.end(), {L.e, L.l, R.e, R.l});
    } else {
      S.faces.insert(S.faces.end(), {L.e, L.l, -1, -1});
    }
    i = j;
  }
  // internal order: block-local grouping by link signature (as for triangles)
  auto sig = [&](int e) {
    int s = 0;
    for (int l = 0; l < 4; l++) s = s * 32 + (nbe[e * 4 + l] >= 0 ? (nbo[e * 4 + l] * 4 + nbl[e * 4 + l]) : 31);
    return s;
  };
  // Locality first: elements along a Morton curve of their centroids, quantised
  // to cells of about one box of 6 tetrahedra (the cube root of 6 |bbox| / E),
  // so a block of 384 elements is a compact brick and face neighbours across
  // blocks sit in nearby blocks (the tiles of a workgroup window share L2 lines
  // instead of each face of a lexicographic order reaching a whole slab away).
  std::vector<uint64_t> key(E);
  {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < (size_t)E * 4; i++)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], m->xyz[i * 3 + d]);
        hi[d] = std::max(hi[d], m->xyz[i * 3 + d]);
      }
    const double vol = std::max(1e-300, (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]));
    const double h = std::cbrt(6.0 * vol / E);
    auto spread = [](uint64_t x) {  // 21 bits -> every third bit
      x &= 0x1fffff;
      x = (x | x << 32) & 0x1f00000000ffffULL;
      x = (x | x << 16) & 0x1f0000ff0000ffULL;
      x = (x | x << 8) & 0x100f00f00f00f00fULL;
      x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
      x = (x | x << 2) & 0x1249249249249249ULL;
      return x;
    };
    for (int e = 0; e < E; e++) {
      uint64_t q[3];
      for (int d = 0; d < 3; d++) {
        double c = 0.0;
        for (int v = 0; v < 4; v++) c += m->xyz[((size_t)e * 4 + v) * 3 + d];
        const double t = (0.25 * c - lo[d]) / h;
        q[d] = (uint64_t)std::min(std::max(t, 0.0), (double)0x1fffff);
      }
      key[e] = spread(q[0]) | spread(q[1]) << 1 | spread(q[2]) << 2;
    }
  }
  std::vector<int32_t> loc(E);
  for (int e = 0; e < E; e++) loc[e] = e;
  std::stable_sort(loc.begin(), loc.end(), [&](int a, int b) { return key[a] < key[b]; });
  // then block-local grouping by link signature (as for triangles)
  std::vector<int> sg(E);
  for (int e = 0; e < E; e++) sg[e] = sig(e);
  const int BLK = 6 * 64;
  S.perm = loc;
  for (int b0 = 0; b0 < E; b0 += BLK)
    std::stable_sort(S.perm.begin() + b0, S.perm.begin() + std::min(E, b0 + BLK),
                     [&](int a, int b) { return sg[a] < sg[b]; });
  std::vector<int32_t> iperm(E);
  for (int i = 0; i < E; i++) iperm[S.perm[i]] = i;
  if ((long)E >= (1L << 26)) return failS(err, DG_ERR_ARG, "too many elements for the link encoding");

  S.mg.assign((size_t)9 * E, 0.0);
  S.detj.assign(E, 0.0);
  S.size.assign(E, 0.0);
  S.link.assign((size_t)4 * E, 0);
  for (int i = 0; i < E; i++) {
    const int e = S.perm[i];
    const double* X = m->xyz + (size_t)e * 12;
    double J[3][3];
    for (int d = 0; d < 3; d++)
      for (int c = 0; c < 3; c++) J[d][c] = X[(c + 1) * 3 + d] - X[d];  // dx_d / dxi_c
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return failS(err, DG_ERR_ARG, "degenerate element %d", e);
    // |J| (J^-1)[r][d] = sign(det) * cofactor(J)[d][r]
    double cof[3][3];
    cof[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    cof[0][1] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);
    cof[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    cof[1][0] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]);
    cof[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    cof[1][2] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]);
    cof[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    cof[2][1] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]);
    cof[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double sgn = det > 0 ? 1.0 : -1.0;
    for (int r = 0; r < 3; r++)
      for (int d = 0; d < 3; d++) S.mg[(size_t)(r * 3 + d) * E + i] = sgn * cof[d][r];
    S.detj[i] = std::fabs(det);
    static const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    double h = 1e300;
    for (int k = 0; k < 6; k++) {
      double s2 = 0;
      for (int d = 0; d < 3; d++) {
        const double v = X[ed[k][1] * 3 + d] - X[ed[k][0] * 3 + d];
        s2 += v * v;
      }
      h = std::min(h, std::sqrt(s2));
    }
    S.size[i] = h;
    for (int l = 0; l < 4; l++) {
      int32_t v;
      if (nbe[e * 4 + l] >= 0)
        v = (iperm[nbe[e * 4 + l]] << 5) | (nbo[e * 4 + l] << 2) | nbl[e * 4 + l];
      else
        v = -BC_PEC;  // every boundary face of a Maxwell mesh is a perfect conductor
      S.link[(size_t)l * E + i] = v;
    }
  }
  S.tab.assign(TetBlock::size(P), 0.0);
  try {
    build_tet_block(P, S.tab.data());
  } catch (const std::exception& ex) {
    return failS(err, DG_ERR_ARG, "%s", ex.what());
  }
  return DG_OK;
}

bool tri_point_field(const std::vector<double>& xyz, const std::vector<int32_t>& perm, int P, int nval,
                     dg_field_fn fn, void* user, std::vector<double>& out) {
  const int E = (int)perm.size(), NQ = nqv_tri(P), NQF = nqf_tri(P), NPT = NQ + 3 * NQF;
  auto vr = tri_rule(2 * P);
  auto gl = gl_rule(NQF);
  std::vector<double> xi((size_t)NPT * 2);
  for (int q = 0; q < NQ; q++) {
    xi[2 * q] = vr[q].x[0];
    xi[2 * q + 1] = vr[q].x[1];
  }
  for (int l = 0; l < 3; l++)
    for (int q = 0; q < NQF; q++) {
      double r[3];
      edge_point(l, gl[q].x[0], r);
      xi[2 * (NQ + l * NQF + q)] = r[0];
      xi[2 * (NQ + l * NQF + q) + 1] = r[1];
    }
  out.assign((size_t)NPT * nval * E, 0.0);
  for (int i = 0; i < E; i++) {
    const double* X = &xyz[(size_t)perm[i] * 6];
    for (int q = 0; q < NPT; q++) {
      const double a = xi[2 * q], b = xi[2 * q + 1];
      const double x[3] = {X[0] + a * (X[2] - X[0]) + b * (X[4] - X[0]), X[1] + a * (X[3] - X[1]) + b * (X[5] - X[1]),
                           0.0};
      double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      fn(x, 0.0, v, user);
      for (int k = 0; k < nval; k++) {
        if (!std::isfinite(v[k])) return false;
        out[((size_t)q * nval + k) * E + i] = v[k];
      }
    }
  }
  return true;
}

}  // namespace dgp
