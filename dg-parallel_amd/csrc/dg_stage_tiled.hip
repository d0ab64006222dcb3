// HIP kernel wrapper of the tiled triangle stage (phases in dg_stage_tiled.h).
#include <hip/hip_runtime.h>

#include "dg_stage_tiled.h"

namespace dgp {

// Workgroups are handed to the 8 XCDs round-robin; give each XCD a contiguous
// run of element tiles so face-neighbour reads stay in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int full = (nb / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

template <int P, class LAW, typename CT, typename IT, typename OT>
__global__ __launch_bounds__(TILE_THREADS) void tri_stage_tiled(StageArgs a) {
  using K = TiledStage<P, LAW, CT, IT, OT>;
  __shared__ typename K::Smem sm;
  typename K::Regs g;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  K::phaseA(a, sm, tid, tile, g);
  __syncthreads();
  K::phaseB(a, sm, tid, g);
  __syncthreads();
  K::phaseD1(a, sm, tid, g);
  __syncthreads();
  K::phaseC1(a, sm, tid, tile, g);
  __syncthreads();
  K::phaseC2(a, sm, tid, g);
  __syncthreads();
  K::phaseD2(a, sm, tid, tile, g);
  if (a.resid_part) {
    // deterministic tree reduction of the step residual norm
    sm.red[tid] = g.rsum;
    __syncthreads();
    for (int s = TILE_THREADS / 2; s > 0; s >>= 1) {
      if (tid < s) sm.red[tid] += sm.red[tid + s];
      __syncthreads();
    }
    if (tid == 0) a.resid_part[tile] = sm.red[0];
  }
}

int tri_tiled_blocks(int E) { return (E + TILE_ELEMS - 1) / TILE_ELEMS; }

template <int P, class LAW>
static int launch_tiled_p(const StageArgs& a, int variant, hipStream_t s) {
  dim3 grid(tri_tiled_blocks(a.E)), block(TILE_THREADS);
  switch (variant) {
    case 0:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, double, double, double>), grid, block, 0, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, double, float>), grid, block, 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, float, float>), grid, block, 0, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, float, double>), grid, block, 0, s, a);
      break;
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class LAW>
static int launch_tiled_law(const StageArgs& a, int variant, hipStream_t s) {
  switch (a.P) {
    case 1: return launch_tiled_p<1, LAW>(a, variant, s);
    case 2: return launch_tiled_p<2, LAW>(a, variant, s);
    case 3: return launch_tiled_p<3, LAW>(a, variant, s);
    case 4: return launch_tiled_p<4, LAW>(a, variant, s);
  }
  return -1;
}

int launch_tri_stage_tiled(const StageArgs& a, int law, int variant, hipStream_t s) {
  if (law == LAW_SCALAR) return launch_tiled_law<ScalarLawD>(a, variant, s);
  if (law == LAW_EULER2D) return launch_tiled_law<EulerLawD>(a, variant, s);
  return -1;
}

}  // namespace dgp
