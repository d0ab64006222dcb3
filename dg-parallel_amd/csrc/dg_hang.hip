// The hanging-face pass (phases in dg_hang.h): one workgroup per element with
// faces of half an edge.
#include <hip/hip_runtime.h>

#include "dg_hang.h"

namespace dgp {

template <int P, class LAW>
__global__ __launch_bounds__(HANG_LANES) void tri_hang_kernel(HangArgs a) {
  using K = HangBlock<P, LAW>;
  __shared__ typename K::Smem sm;
  const int i = blockIdx.x, lane = threadIdx.x;
  K::hangA(a, i, lane, sm);
  __syncthreads();
  K::hangB(a, i, lane, sm);
  __syncthreads();
  K::hangC(a, i, lane, sm);
}

template <class LAW>
static int launch_hang_law(const HangArgs& a, hipStream_t s) {
  dim3 grid(a.n), block(HANG_LANES);
  switch (a.P) {
    case 1: hipLaunchKernelGGL((tri_hang_kernel<1, LAW>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((tri_hang_kernel<2, LAW>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((tri_hang_kernel<3, LAW>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((tri_hang_kernel<4, LAW>), grid, block, 0, s, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_hang(const HangArgs& a, int law, hipStream_t s) {
  if (a.n <= 0) return 0;
  if (law == LAW_SCALAR) return launch_hang_law<ScalarLawD>(a, s);
  if (law == LAW_EULER2D && a.law.flux == FLUX_EXACT) return launch_hang_law<EulerExactLawD>(a, s);
  if (law == LAW_EULER2D) return launch_hang_law<EulerLawD>(a, s);
  if (law == LAW_MAXWELL2D) return launch_hang_law<Maxwell2dLawD>(a, s);
  if (law == LAW_BURGERS2D) return launch_hang_law<Burgers2DLawD>(a, s);
  return -1;
}

}  // namespace dgp
