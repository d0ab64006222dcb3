// Tiled triangle stage kernel, tile shape NW = 4 (256 threads per 64 elements),
// and the launcher over every tile shape.
#include "dg_stage_tiled_kern.h"

namespace dgp {

template int launch_tiled_nw<4>(const StageArgs&, int, int, hipStream_t);

int tri_tiled_blocks(int E) { return (E + TILE_ELEMS - 1) / TILE_ELEMS; }

// variant: bits 0-1 the precision form (0 FP64; 1 FP32 arithmetic with FP64
// stage input; 2 FP32 throughout with FP64 accumulation; 3 FP32 with an FP64
// output), bits 2-3 the tile shape (0: 4 waves per tile, 1: 8 waves, 2: 4 waves
// with the contractions on the matrix cores)
int launch_tri_stage_tiled(const StageArgs& a, int law, int variant, hipStream_t s) {
  const int prec = variant & 3;
  switch (variant >> 2) {
    case 0: return launch_tiled_nw<4>(a, law, prec, s);
    case 1: return launch_tiled_nw<8>(a, law, prec, s);
    case 2: return launch_tiled_mfma(a, law, prec, s);
  }
  return -1;
}

}  // namespace dgp
