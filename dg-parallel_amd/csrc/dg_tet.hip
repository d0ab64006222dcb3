// HIP kernel of the tetrahedron stage, FP64 form (phases in dg_tet.h, kernel
// and launcher in dg_tet_kern.h); dispatch of every precision form.
#include "dg_tet_kern.h"

namespace dgp {

int tet_tiles(int E) { return (E + TILE_ELEMS - 1) / TILE_ELEMS; }

int launch_tet_stage(const TetArgs& a, int shape, int prec, hipStream_t s) {
  if (prec == 0) return launch_tet_shape<double, double, double>(a, shape, s);
  return launch_tet_stage32(a, shape, prec, s);
}

}  // namespace dgp
