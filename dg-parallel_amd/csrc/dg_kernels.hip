// Hand-written gfx950 kernels for the explicit DG residual and Runge-Kutta
// stage on triangles: the element-per-work-item stage kernel (element code in
// dg_element.h), the time step, reductions and probes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "dg_element.h"
#include "dg_kernels.h"
#include "dg_tables.h"

namespace dgp {

constexpr int BLOCK = 256;

// Workgroups are handed to the 8 XCDs round-robin; give each XCD a contiguous
// run of element tiles so face-neighbour reads hit that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  int full = (nb / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

template <int P, class LAW, typename CT, typename IT, typename OT>
__global__ __launch_bounds__(BLOCK) void tri_stage_kernel(StageArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int e = blk * BLOCK + threadIdx.x;
  double rsum = 0.0;
  if (e < a.E) element_stage<P, LAW, CT, IT, OT>(a, e, rsum);
  if (a.resid_part) {
    // deterministic block reduction of the residual norm
    __shared__ double red[BLOCK];
    red[threadIdx.x] = rsum;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.resid_part[blk] = red[0];
  }
}

int tri_stage_blocks(int E) { return (E + BLOCK - 1) / BLOCK; }

template <int P, class LAW>
static int launch_p(const StageArgs& a, int variant, hipStream_t s) {
  dim3 grid(tri_stage_blocks(a.E)), block(BLOCK);
  switch (variant) {
    case 0:
      hipLaunchKernelGGL((tri_stage_kernel<P, LAW, double, double, double>), grid, block, 0, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((tri_stage_kernel<P, LAW, float, double, float>), grid, block, 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((tri_stage_kernel<P, LAW, float, float, float>), grid, block, 0, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((tri_stage_kernel<P, LAW, float, float, double>), grid, block, 0, s, a);
      break;
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class LAW>
static int launch_law(const StageArgs& a, int variant, hipStream_t s) {
  switch (a.P) {
    case 1: return launch_p<1, LAW>(a, variant, s);
    case 2: return launch_p<2, LAW>(a, variant, s);
    case 3: return launch_p<3, LAW>(a, variant, s);
    case 4: return launch_p<4, LAW>(a, variant, s);
  }
  return -1;
}

int launch_tri_stage(const StageArgs& a, int law, int variant, hipStream_t s) {
  if (law == LAW_SCALAR) return launch_law<ScalarLawD>(a, variant, s);
  if (law == LAW_EULER2D && a.law.flux == FLUX_EXACT) return launch_law<EulerExactLawD>(a, variant, s);
  if (law == LAW_EULER2D) return launch_law<EulerLawD>(a, variant, s);

  return -1;
}

// ---------------------------------------------------------------------------
// time-step candidates (DGCell.cc:1155-1207): lambda = max over the mean and the
// three vertex values; dt_e = CFL size_e / ((2p+1) lambda); per-block minima.
template <int P, class LAW>
__global__ __launch_bounds__(BLOCK) void tri_dt_kernel(DtArgs a) {
  constexpr int NP = np_tri(P), NF = LAW::NF;
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  double dt = 1e300;
  if (e < a.E) {
    double U[NF][NP];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int k = 0; k < NP; k++) U[f][k] = a.U[(long)(f * NP + k) * a.E + e];
    double mean[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) mean[f] = U[f][0] * a.vtx[0];
    double lam;
    if constexpr (point_vel<LAW>::value) {
      // ScalarLaw::maximumEigenValue at the cell's first volume point (DGCell.cc:1165-1186)
      lam = fabs(a.pvel[e]) + fabs(a.pvel[(long)a.E + e]);
    } else {
      lam = LAW::maxev(a.law, mean);
#pragma unroll
      for (int v = 0; v < 3; v++) {
        double val[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          double s = 0;
#pragma unroll
          for (int k = 0; k < NP; k++) s += U[f][k] * a.vtx[(1 + v) * NP + k];
          val[f] = s;
        }
        double b = LAW::maxev(a.law, val);
        lam = b > lam ? b : lam;
      }
    }
    dt = a.el.size[e] * a.cfl / (lam * (2. * P + 1.0));
    if (a.dtscale) dt *= a.dtscale[e];
  }
  __shared__ double red[BLOCK];
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}

int tri_dt_blocks(int E) { return (E + BLOCK - 1) / BLOCK; }

template <class LAW>
static int launch_dt_law(const DtArgs& a, hipStream_t s) {
  dim3 grid(tri_dt_blocks(a.E)), block(BLOCK);
  switch (a.P) {
    case 1: hipLaunchKernelGGL((tri_dt_kernel<1, LAW>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((tri_dt_kernel<2, LAW>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((tri_dt_kernel<3, LAW>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((tri_dt_kernel<4, LAW>), grid, block, 0, s, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_tri_dt(const DtArgs& a, int law, hipStream_t s) {
  if (law == LAW_SCALAR && a.pvel) return launch_dt_law<ScalarFieldLawD>(a, s);
  if (law == LAW_SCALAR) return launch_dt_law<ScalarLawD>(a, s);
  if (law == LAW_EULER2D) return launch_dt_law<EulerLawD>(a, s);
  if (law == LAW_MAXWELL2D) return launch_dt_law<Maxwell2dLawD>(a, s);
  if (law == LAW_BURGERS2D) return launch_dt_law<Burgers2DLawD>(a, s);
  return -1;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free time step: the per-workgroup minima of the dt
// kernels.  Host code never passes `scratch` itself to a kernel (on the host it
// names a shadow variable, not device memory); it passes the address that
// hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static double* scratch_addr[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return -1;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_addr[device]) return 0;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    (void)hipGetLastError();
    return -2;
  }
  scratch_addr[device] = (double*)p;
  return 0;
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(scratch_mu);
  return scratch_addr[device];
}

__global__ __launch_bounds__(BLOCK) void min_reduce_kernel(const double* __restrict__ part, int n,
                                                           double* __restrict__ out) {
  double v = 1e300;
  for (int i = threadIdx.x; i < n; i += BLOCK) v = fmin(v, part[i]);
  __shared__ double red[BLOCK];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void fill_one_kernel(double* out, double v) { out[0] = v; }

__global__ __launch_bounds__(BLOCK) void nonfinite_scan_kernel(const double* __restrict__ U, long n,
                                                               unsigned* __restrict__ status) {
  bool bad = false;
  for (long i = blockIdx.x * (long)BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
    const double x = U[i];
    bad |= !(x - x == 0.0);
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 4u);
}

int launch_nonfinite_scan(const double* U, long n, unsigned* status, hipStream_t s) {
  const long blocks = std::min<long>((n + BLOCK - 1) / BLOCK, 4096);
  hipLaunchKernelGGL(nonfinite_scan_kernel, dim3((unsigned)std::max<long>(blocks, 1)), dim3(BLOCK), 0, s, U, n,
                     status);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_min_reduce(const double* part, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(min_reduce_kernel, dim3(1), dim3(BLOCK), 0, s, part, n, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_fill_one(double* out, double v, hipStream_t s) {
  hipLaunchKernelGGL(fill_one_kernel, dim3(1), dim3(1), 0, s, out, v);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}


}  // namespace dgp

namespace dgp {

// plain streaming copy: the reference point for HBM byte counters
__global__ __launch_bounds__(256) void copy_probe_kernel(const double4* __restrict__ src, double4* __restrict__ dst,
                                                          long n4) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n4; i += (long)gridDim.x * 256) dst[i] = src[i];
}

// out[i] = in[i] rounded to FP32 (the FP32 form's first-stage input)
__global__ __launch_bounds__(256) void round_f32_kernel(const double* __restrict__ in, float* __restrict__ out, long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) out[i] = float(in[i]);
}

int launch_round_f32(const double* in, float* out, long n, hipStream_t s) {
  const long nb = std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(round_f32_kernel, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, in, out, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_copy_probe(const void* src, void* dst, long n_doubles, hipStream_t s) {
  long n4 = n_doubles / 4;
  hipLaunchKernelGGL(copy_probe_kernel, dim3(4096), dim3(256), 0, s, (const double4*)src, (double4*)dst, n4);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace dgp
