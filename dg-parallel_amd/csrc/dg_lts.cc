// Local time stepping: levels, element classes and the launch schedule of one
// step (the method is stated in dg_lts.h).
#include "dg_lts.h"

#include <algorithm>
#include <tuple>

namespace dgp {

int lts_levels(int E, const double* size, const std::vector<int32_t>& faces, int max_levels,
               std::vector<int32_t>& level) {
  if (max_levels < 1) max_levels = 1;
  if (max_levels > LTS_MAX_LEVELS) max_levels = LTS_MAX_LEVELS;
  double smax = 0.0;
  for (int e = 0; e < E; e++) smax = std::max(smax, size[e]);
  level.assign(E, 0);
  for (int e = 0; e < E; e++) {
    // the bins' upper bounds BB[b] * 1.0001 with BB[b] = smax / 2^b
    int b = 0;
    double bound = smax * 1.0001;
    while (b + 1 < max_levels && size[e] <= 0.5 * bound) {
      bound *= 0.5;
      b++;
    }
    level[e] = b;
  }
  // face neighbours at most one level apart: raise the coarser side
  const size_t nf = faces.size() / 4;
  for (bool changed = true; changed;) {
    changed = false;
    for (size_t f = 0; f < nf; f++) {
      const int a = faces[4 * f], c = faces[4 * f + 2];
      if (c < 0) continue;
      if (level[a] < level[c] - 1) {
        level[a] = level[c] - 1;
        changed = true;
      } else if (level[c] < level[a] - 1) {
        level[c] = level[a] - 1;
        changed = true;
      }
    }
  }
  int top = 0;
  for (int e = 0; e < E; e++) top = std::max(top, level[e]);
  return top + 1;
}

void lts_reach(int E, const std::vector<int32_t>& faces, const std::vector<int32_t>& level,
               std::vector<int32_t>& m1, std::vector<int32_t>& m2) {
  const size_t nf = faces.size() / 4;
  m1 = level;
  for (size_t f = 0; f < nf; f++) {
    const int a = faces[4 * f], c = faces[4 * f + 2];
    if (c < 0) continue;
    m1[a] = std::max(m1[a], level[c]);
    m1[c] = std::max(m1[c], level[a]);
  }
  m2 = m1;
  for (size_t f = 0; f < nf; f++) {
    const int a = faces[4 * f], c = faces[4 * f + 2];
    if (c < 0) continue;
    m2[a] = std::max(m2[a], m1[c]);
    m2[c] = std::max(m2[c], m1[a]);
  }
  (void)E;
}

void lts_classify(int E, const std::vector<int32_t>& level, const std::vector<int32_t>& m1,
                  const std::vector<int32_t>& m2, std::vector<int32_t>& cid, std::vector<LtsClass>& cls) {
  std::vector<std::tuple<int, int, int>> keys;
  for (int e = 0; e < E; e++) keys.emplace_back(level[e], m1[e], m2[e]);
  std::vector<std::tuple<int, int, int>> uniq = keys;
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  cls.clear();
  for (const auto& k : uniq) cls.push_back({std::get<0>(k), std::get<1>(k), std::get<2>(k), 0, 0});
  cid.assign(E, 0);
  for (int e = 0; e < E; e++)
    cid[e] = (int32_t)(std::lower_bound(uniq.begin(), uniq.end(), keys[e]) - uniq.begin());
}

static void push_tiles(const LtsClass& c, int flags, double w, std::vector<LtsTile>& out) {
  for (int e = c.e0; e < c.e1; e += 64) out.push_back({e, std::min(c.e1, e + 64), flags, c.b, w});
}

void lts_plan(int nlev, const std::vector<LtsClass>& cls, LtsPlan& plan) {
  const int S = 1 << nlev;
  plan.S = S;
  plan.tiles.clear();
  plan.ctiles.clear();
  plan.off.assign(1, 0);
  plan.coff.assign(1, 0);
  for (int i = 0; i < S; i++) {
    const bool odd = i & 1;
    for (const LtsClass& c : cls) {
      if (c.e1 <= c.e0) continue;
      const int Lb = S >> c.b, L1 = S >> c.m1, L2 = S >> c.m2, j = i % Lb;
      if (!odd && i % L1 == 0)
        push_tiles(c, j == 0 ? LTS_RESET : 0, (double)L1 / (2.0 * Lb), plan.tiles);
      else if (odd && (i - 1) % L2 == 0)
        // with L2 = Lb the span's only evaluated odd stage is its second one, and
        // no element reads this element's y_m at an even stage before the span
        // ends (that would need a finer element within two faces): the span's
        // increment is complete there and goes to y_m at once
        push_tiles(c, LTS_ODD | (j == Lb - 1 || L2 == Lb ? LTS_END : 0), (double)L2 / (2.0 * Lb), plan.tiles);
      // a span that ends at a stage its elements skip
      if (odd && j == Lb - 1 && (i - 1) % L2 != 0 && L2 != Lb) push_tiles(c, 0, 0.0, plan.ctiles);
    }
    plan.off.push_back((int)plan.tiles.size());
    plan.coff.push_back((int)plan.ctiles.size());
  }
}

}  // namespace dgp
