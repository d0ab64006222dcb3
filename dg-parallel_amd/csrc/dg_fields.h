// Exact fields of the benchmark problems (host FP64).
#pragma once
#include "dg_laws.h"

namespace dgp {

enum { FIELD_GAUSSIAN = 0, FIELD_VORTEX = 1, FIELD_SOD = 2, FIELD_CAVITY = 3 };

// val receives Nf conserved values at point x, time t; returns 0 or -1 (unknown field)
int exact_field(int field, const double* par, const LawParams& law, const double* x, double t, double* val);

}  // namespace dgp
