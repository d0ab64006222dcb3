// The tetrahedron stage kernel (phases in dg_tet.h) and its launcher over tile
// shapes and orders, for one precision form per translation unit (FP64 in
// dg_tet.hip, the FP32 form in dg_tet32.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "dg_tet.h"

namespace dgp {

__device__ __forceinline__ int xcd_tile_tet(int b, int nb) {
  int full = (nb / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

template <int P, int NW, bool ALIAS, typename CT, typename IT, typename OT>
__global__ __launch_bounds__(NW * 64) void tet_stage_tiled(TetArgs a) {
  using K = TiledTet<P, NW, ALIAS, CT, IT, OT>;
  __shared__ typename K::Smem sm;
  typename K::Regs g;
  const int tile = xcd_tile_tet(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  K::phaseA(a, sm, tid, tile);
  __syncthreads();
  K::phaseB(a, sm, tid, tile, g);
#pragma unroll 1
  for (int l = 0; l < 4; l++) {
    __syncthreads();
    K::phaseC1(a, sm, tid, tile, l);
    __syncthreads();
    K::phaseC2(a, sm, tid, tile, l);
    __syncthreads();
    K::phaseLift(a, sm, tid, l, g);
  }
  K::phaseD(a, tid, tile, g);
  if (a.resid_part) {
    __syncthreads();
    sm.red[tid] = g.rsum;
    __syncthreads();
    for (int s = K::THREADS / 2; s > 0; s >>= 1) {
      if (tid < s) sm.red[tid] += sm.red[tid + s];
      __syncthreads();
    }
    if (tid == 0) a.resid_part[tile] = sm.red[0];
  }
}

template <int NW, bool ALIAS, typename CT, typename IT, typename OT>
static int launch_tet_form(const TetArgs& a, hipStream_t s) {
  dim3 grid(tet_tiles(a.E)), block(NW * 64);
  switch (a.P) {
    case 1: hipLaunchKernelGGL((tet_stage_tiled<1, NW, ALIAS, CT, IT, OT>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((tet_stage_tiled<2, NW, ALIAS, CT, IT, OT>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((tet_stage_tiled<3, NW, ALIAS, CT, IT, OT>), grid, block, 0, s, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <typename CT, typename IT, typename OT>
static int launch_tet_shape(const TetArgs& a, int shape, hipStream_t s) {
  switch (shape) {
    case 0: return launch_tet_form<4, true, CT, IT, OT>(a, s);
    case 1: return launch_tet_form<8, true, CT, IT, OT>(a, s);
    case 2: return launch_tet_form<8, false, CT, IT, OT>(a, s);
    case 3: return launch_tet_form<16, false, CT, IT, OT>(a, s);
    case 4: return launch_tet_form<4, false, CT, IT, OT>(a, s);
  }
  return -1;
}

}  // namespace dgp
