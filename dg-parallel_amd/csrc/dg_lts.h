// Local time stepping on triangles: host planning (levels, element classes and
// the launch schedule of one step).  New design; the reference's shape is
// RungeKuttaTVD2adaptive (TimeIntegrator.cc:442-833) over the size bins of
// DGAnalysis::sortCellsBySize (DGAnalysis.cc:749-890), whose range assembly
// overloads are empty (TimeIntegrator.cc:1438-1487), so it advances nothing.
//
// The method.  A step of size h on levels 0..nlev-1: a level-b element takes
// 2^b SSP-RK2 (Heun) substeps of h_b = h / 2^b.  All levels share one stage
// sequence of S = 2^nlev stages.  Level b splits it into 2^b spans of
// L_b = S / 2^b stages; in span m, stage j of the span gives the element the
// stage state
//     Y_j = y_m                 (j even)
//     Y_j = y_m + h_b k_{j-1}   (j odd)
// with k_j = R(Y_j) / |J| the residual of stage j (every element evaluated with
// its neighbours' stage states of the same stage), and at the span's end
//     y_{m+1} = y_m + (h_b / L_b) sum_j k_j.
// This is a partitioned Runge-Kutta method whose partitions (levels) share the
// weights b_i = 1/S, so the face fluxes of every stage enter both sides of a
// face with the same weight: the method is conservative.  For each level,
// sum_i b_i c_i = 1/2, so it is second order; on one level (S = 2) it is SSP-RK2.
//
// Saving work.  Stage j's inputs repeat those of stage j - 2 unless a
// neighbour's state moved between them.  An element's even stages read its own
// y_m and its face neighbours' y_m; they change within a span only through a
// finer element within one face (m1 = the finest level among the element and
// its neighbours): k repeats with period L1 = S / 2^m1.  Its odd stages also
// read the neighbours' predictors, which change with the neighbours' even k:
// period L2 = S / 2^m2, m2 the finest level within two faces.  An element
// therefore evaluates its residual only at even stages i = 0 mod L1 and odd
// stages i = 1 mod L2; a skipped stage's k equals the last evaluated k of the
// same parity exactly (same inputs, same arithmetic), and enters the sum with
// multiplicity L1/2 (even) or L2/2 (odd).  Far from finer elements that is one
// Heun substep per h_b, as on a uniform mesh of that level.  Such an element
// (L2 = L_b) writes its new y_m at the span's second stage: until the span
// ends, only odd stages read it, and they read its predictor.
#pragma once
#include <stdint.h>

#include <vector>

namespace dgp {

constexpr int LTS_MAX_LEVELS = 6;

// Levels by cell size, as DGAnalysis::sortCellsBySize bins them (DGAnalysis.cc:
// 749-803): the largest b < max_levels with size <= 1.0001 smax / 2^b.  Then
// levels are raised until face neighbours differ by at most one level (the
// reference stops with "Too big a difference between elements" instead,
// DGAnalysis.cc:836); a finer level only shortens an element's step.
// faces: (eL, lL, eR, lR), eR < 0 on boundary faces.  Returns the number of
// levels in use (max level + 1).
int lts_levels(int E, const double* size, const std::vector<int32_t>& faces, int max_levels,
               std::vector<int32_t>& level);

// m1[e], m2[e]: the finest level within one and within two faces of e
void lts_reach(int E, const std::vector<int32_t>& faces, const std::vector<int32_t>& level,
               std::vector<int32_t>& m1, std::vector<int32_t>& m2);

// an element class: level b, reach m1 <= m2, and its internal element range
struct LtsClass {
  int b, m1, m2;
  int e0, e1;
};

// class id of every element (classes ordered by (b, m1, m2)); cls gets the
// classes with empty ranges
void lts_classify(int E, const std::vector<int32_t>& level, const std::vector<int32_t>& m1,
                  const std::vector<int32_t>& m2, std::vector<int32_t>& cid, std::vector<LtsClass>& cls);

// stage flags: bit 0 odd stage, bit 1 first stage of a span (the increment
// restarts), bit 2 the span ends here (the new y_m is written)
enum { LTS_ODD = 1, LTS_RESET = 2, LTS_END = 4 };

// One tile (at most 64 elements of one class) of a launch: elements [e0, e1),
// the class's stage flags, its level b (substep h / 2^b) and the weight w of
// the stage's h_b k in the span increment.
struct LtsTile {
  int e0, e1, flags, b;
  double w;
};

// The launches of one step: stage i evaluates the residual of the tiles
// tiles[off[i] .. off[i+1]) on the stage input (even stages: the span-start
// states y_m; odd stages: the predictors) and updates them; then, at odd
// stages, the tiles ctiles[coff[i] .. coff[i+1]) (classes whose span ends at a
// stage they skip) add the span's increment to y_m.  One launch each, in this
// order; the plan does not depend on the step size.
struct LtsPlan {
  int S = 0;
  std::vector<LtsTile> tiles, ctiles;
  std::vector<int> off, coff;
};
void lts_plan(int nlev, const std::vector<LtsClass>& cls, LtsPlan& plan);

}  // namespace dgp
