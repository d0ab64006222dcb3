// Tiled triangle stage kernel, tile shape NW = 8 (512 threads per 64 elements).
#include "dg_stage_tiled_kern.h"

namespace dgp {

template int launch_tiled_nw<8>(const StageArgs&, int, int, hipStream_t);

}  // namespace dgp
