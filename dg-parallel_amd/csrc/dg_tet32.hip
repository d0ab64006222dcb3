// HIP kernel of the tetrahedron stage, FP32 form: FP32 arithmetic and tables,
// FP32 stage copies, FP64 step-start state, accumulator and update (phases in
// dg_tet.h, kernel and launcher in dg_tet_kern.h).
#include "dg_tet_kern.h"

namespace dgp {

int launch_tet_stage32(const TetArgs& a, int shape, int prec, hipStream_t s) {
  switch (prec) {
    case 1: return launch_tet_shape<float, double, float>(a, shape, s);
    case 2: return launch_tet_shape<float, float, float>(a, shape, s);
    case 3: return launch_tet_shape<float, float, double>(a, shape, s);
  }
  return -1;
}

}  // namespace dgp
