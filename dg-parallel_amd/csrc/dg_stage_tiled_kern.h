// The tiled triangle stage kernel (phases in dg_stage_tiled.h) and its launcher
// for one tile shape NW (waves per 64-element tile); each shape's
// instantiations live in a translation unit of their own.
#pragma once

#include <hip/hip_runtime.h>

#include "dg_stage_tiled.h"

namespace dgp {

// blockIdx -> tile: consecutive tiles go to one XCD (workgroups are dealt to the
// 8 XCDs round robin), so neighbouring tiles share that XCD's L2
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int full = (nb / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

// the phases of one tile, separated by workgroup barriers
template <class K, bool MF>
__device__ __forceinline__ void tiled_body(const StageArgs& a, typename K::Smem& sm, int tile) {
  typename K::Regs g;
  const int tid = threadIdx.x;
  K::phaseA(a, sm, tid, tile, g);
  __syncthreads();
#pragma unroll
  for (int q0 = 0; q0 < K::NQ; q0 += K::QH) {
    K::phaseB(a, sm, tid, g, q0);
    __syncthreads();
    K::phaseD1(a, sm, tid, g, q0);
    __syncthreads();
  }
  K::phaseC1(a, sm, tid, tile, g);
  __syncthreads();
  K::phaseC2(a, sm, tid, g);
  __syncthreads();
  if constexpr (MF) {
    K::phaseLiftMF(a, sm, tid, g);
    __syncthreads();  // every wave has read the face fluxes
    K::phaseStoreMF(sm, tid, g);
    __syncthreads();
  }
  K::phaseD2(a, sm, tid, tile, g);
  if (a.resid_part) {
    // deterministic tree reduction of the step residual norm (red overlays the
    // face fluxes D2 reads)
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

template <int P, class LAW, typename CT, typename IT, typename OT, int NW, bool MF = false>
__global__ __launch_bounds__(NW * 64) void tri_stage_tiled(StageArgs a) {
  using K = TiledStage<P, LAW, CT, IT, OT, NW, MF>;
  __shared__ typename K::Smem sm;
  tiled_body<K, MF>(a, sm, xcd_tile(blockIdx.x, gridDim.x));
}

// tiles of the launch's element range [e0, e1) (e1 = 0: the whole mesh)
static inline int tiled_grid(const StageArgs& a) {
  return ((a.e1 > 0 ? a.e1 : a.E) - a.e0 + TILE_ELEMS - 1) / TILE_ELEMS;
}

template <int P, class LAW, int NW>
static int launch_tiled_p(const StageArgs& a, int prec, hipStream_t s) {
  dim3 grid(tiled_grid(a)), block(NW * 64);
  switch (prec) {
    case 0:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, double, double, double, NW>), grid, block, 0, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, double, float, NW>), grid, block, 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, float, float, NW>), grid, block, 0, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, float, double, NW>), grid, block, 0, s, a);
      break;
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class LAW, int NW>
static int launch_tiled_law(const StageArgs& a, int prec, hipStream_t s) {
  switch (a.P) {
    case 1: return launch_tiled_p<1, LAW, NW>(a, prec, s);
    case 2: return launch_tiled_p<2, LAW, NW>(a, prec, s);
    case 3: return launch_tiled_p<3, LAW, NW>(a, prec, s);
    case 4: return launch_tiled_p<4, LAW, NW>(a, prec, s);
  }
  return -1;
}

// the MFMA tile shape: FP32 forms (prec 1..3) of the four-field law
template <int P, class LAW>
static int launch_tiled_mf_p(const StageArgs& a, int prec, hipStream_t s) {
  dim3 grid(tiled_grid(a)), block(4 * 64);
  switch (prec) {
    case 1:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, double, float, 4, true>), grid, block, 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, float, float, 4, true>), grid, block, 0, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((tri_stage_tiled<P, LAW, float, float, double, 4, true>), grid, block, 0, s, a);
      break;
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class LAW>
static int launch_tiled_mf_law(const StageArgs& a, int prec, hipStream_t s) {
  switch (a.P) {
    case 1: return launch_tiled_mf_p<1, LAW>(a, prec, s);
    case 2: return launch_tiled_mf_p<2, LAW>(a, prec, s);
    case 3: return launch_tiled_mf_p<3, LAW>(a, prec, s);
    case 4: return launch_tiled_mf_p<4, LAW>(a, prec, s);
  }
  return -1;
}

template <int NW>
int launch_tiled_nw(const StageArgs& a, int law, int prec, hipStream_t s) {
  if (law == LAW_SCALAR && a.pvel) return launch_tiled_law<ScalarFieldLawD, NW>(a, prec, s);
  if (law == LAW_SCALAR) return launch_tiled_law<ScalarLawD, NW>(a, prec, s);
  if (law == LAW_EULER2D && a.law.char_bc)
    return a.law.flux == FLUX_EXACT ? launch_tiled_law<EulerExactCharLawD, NW>(a, prec, s)
                                    : launch_tiled_law<EulerCharLawD, NW>(a, prec, s);
  if (law == LAW_EULER2D && a.law.flux == FLUX_EXACT) return launch_tiled_law<EulerExactLawD, NW>(a, prec, s);
  if (law == LAW_EULER2D) return launch_tiled_law<EulerLawD, NW>(a, prec, s);
  if (law == LAW_MAXWELL2D) return launch_tiled_law<Maxwell2dLawD, NW>(a, prec, s);
  if (law == LAW_BURGERS2D) return launch_tiled_law<Burgers2DLawD, NW>(a, prec, s);
  return -1;
}

extern template int launch_tiled_nw<4>(const StageArgs&, int, int, hipStream_t);
// the MFMA tile shape (dg_stage_tiled_mf.hip): the Euler law's FP32 forms
int launch_tiled_mfma(const StageArgs& a, int law, int prec, hipStream_t s);
extern template int launch_tiled_nw<8>(const StageArgs&, int, int, hipStream_t);

}  // namespace dgp
