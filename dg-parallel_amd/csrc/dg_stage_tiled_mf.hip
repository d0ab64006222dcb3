// Tiled triangle stage kernel, MFMA tile shape: the FP32 forms of the Euler law
// with the test contraction and the face lift on the matrix cores.
#include "dg_stage_tiled_kern.h"

namespace dgp {

int launch_tiled_mfma(const StageArgs& a, int law, int prec, hipStream_t s) {
  if (law != LAW_EULER2D || prec < 1 || prec > 3 || a.law.char_bc) return -1;
  if (a.law.flux == FLUX_EXACT) return launch_tiled_mf_law<EulerExactLawD>(a, prec, s);
  return launch_tiled_mf_law<EulerLawD>(a, prec, s);
}

}  // namespace dgp
