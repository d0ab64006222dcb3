// HIP kernels of the residual derivatives (phases in dg_adjoint.h).
#include <hip/hip_runtime.h>

#include "dg_adjoint.h"

namespace dgp {

__device__ __forceinline__ int xcd_tile_adj(int b, int nb) {
  int full = (nb / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

template <int P, class LAW, int ND>
__global__ __launch_bounds__(TILE_THREADS) void tri_vjp_tiled(AdjArgs a) {
  using K = TiledVjp<P, LAW, ND>;
  __shared__ typename K::Smem sm;
  typename K::Regs g;
  const int tile = xcd_tile_adj(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  K::phaseA(a, sm, tid, tile);
  __syncthreads();
  K::phaseB(a, sm, tid);
  __syncthreads();
  K::phaseD1(a, sm, tid, g);
#pragma unroll 1
  for (int l = 0; l < 3; l++) {
    __syncthreads();
    K::phaseC1(a, sm, tid, tile, l);
    __syncthreads();
    K::phaseC2(a, sm, tid, l);
    __syncthreads();
    K::phaseLift(a, sm, tid, l, g);
  }
  K::phaseOut(a, tid, tile, g);
}

template <int P, class LAW>
__global__ __launch_bounds__(256) void tri_jvp(AdjArgs a) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < a.E) jvp_element<P, LAW>(a, e);
}

// out = (c dt / |J_e|) in, per element over all its rows
__global__ __launch_bounds__(256) void tri_adj_scale(int E, int nrow, double c, double dt, const double* detj,
                                                     const double* in, double* out) {
  const long n = (long)E * nrow;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int e = (int)(i % E);
    out[i] = c * dt / detj[e] * in[i];
  }
}

template <class LAW, int ND>
static int launch_vjp_nd(const AdjArgs& a, hipStream_t s) {
  dim3 grid((a.E + TILE_ELEMS - 1) / TILE_ELEMS), block(TILE_THREADS);
  switch (a.P) {
    case 1: hipLaunchKernelGGL((tri_vjp_tiled<1, LAW, ND>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((tri_vjp_tiled<2, LAW, ND>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((tri_vjp_tiled<3, LAW, ND>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((tri_vjp_tiled<4, LAW, ND>), grid, block, 0, s, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// nd: flux tangent directions per pass.  The default (0) is one: the Euler flux
// Jacobian built in 4 single-direction passes keeps the kernel at 254 VGPRs (two
// waves per SIMD) and ran 1.4x faster than 4 directions at once (440 registers,
// one wave) on the P3 vortex.
int launch_tri_vjp(const AdjArgs& a, int law, int nd, hipStream_t s) {
  if (law == LAW_SCALAR) return launch_vjp_nd<ScalarLawD, 1>(a, s);
  if (law == LAW_BURGERS2D) return launch_vjp_nd<Burgers2DLawD, 1>(a, s);
  if (law == LAW_MAXWELL2D) return launch_vjp_nd<Maxwell2dLawD, 1>(a, s);
  if (law == LAW_EULER2D && a.law.char_bc) {
    if (a.law.flux == FLUX_EXACT) return launch_vjp_nd<EulerExactCharLawD, 1>(a, s);
    return launch_vjp_nd<EulerCharLawD, 1>(a, s);
  }
  if (law == LAW_EULER2D && a.law.flux == FLUX_EXACT) {
    switch (nd) {
      case 0:
      case 1: return launch_vjp_nd<EulerExactLawD, 1>(a, s);
      case 2: return launch_vjp_nd<EulerExactLawD, 2>(a, s);
      case 4: return launch_vjp_nd<EulerExactLawD, 4>(a, s);
    }
    return -1;
  }
  if (law == LAW_EULER2D) {
    switch (nd) {
      case 0:
      case 1: return launch_vjp_nd<EulerLawD, 1>(a, s);
      case 2: return launch_vjp_nd<EulerLawD, 2>(a, s);
      case 4: return launch_vjp_nd<EulerLawD, 4>(a, s);
    }
  }
  return -1;
}

template <class LAW>
static int launch_jvp_law(const AdjArgs& a, hipStream_t s) {
  dim3 grid((a.E + 255) / 256), block(256);
  switch (a.P) {
    case 1: hipLaunchKernelGGL((tri_jvp<1, LAW>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((tri_jvp<2, LAW>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((tri_jvp<3, LAW>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((tri_jvp<4, LAW>), grid, block, 0, s, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_tri_jvp(const AdjArgs& a, int law, hipStream_t s) {
  if (law == LAW_SCALAR) return launch_jvp_law<ScalarLawD>(a, s);
  if (law == LAW_EULER2D && a.law.char_bc)
    return a.law.flux == FLUX_EXACT ? launch_jvp_law<EulerExactCharLawD>(a, s) : launch_jvp_law<EulerCharLawD>(a, s);
  if (law == LAW_EULER2D && a.law.flux == FLUX_EXACT) return launch_jvp_law<EulerExactLawD>(a, s);
  if (law == LAW_EULER2D) return launch_jvp_law<EulerLawD>(a, s);
  if (law == LAW_MAXWELL2D) return launch_jvp_law<Maxwell2dLawD>(a, s);
  if (law == LAW_BURGERS2D) return launch_jvp_law<Burgers2DLawD>(a, s);
  return -1;
}

int launch_tri_adj_scale(int E, int nrow, double c, double dt, const double* detj, const double* in, double* out,
                         hipStream_t s) {
  long n = (long)E * nrow;
  int blocks = (int)std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(tri_adj_scale, dim3(blocks), dim3(256), 0, s, E, nrow, c, dt, detj, in, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace dgp
