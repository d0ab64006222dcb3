// HIP kernels of the mixed-order path (tile phases in dg_mixed_tiled.h, element
// functions in dg_mixed.h).
//
// tri_mixed_stage is one persistent launch for every order bucket: the host
// splits each bucket into chunks of MIX_CHUNK = 64 elements (one tile), heaviest
// order first, and each workgroup walks the chunk list with a stride of the
// grid size.  A chunk never spans buckets, so the order switch is uniform per
// workgroup.  All workgroups reach the end of the list; the residual norm is
// reduced per workgroup in a fixed order.  tri_mixed_stage_p<LAW, P> is the
// other form: one launch per bucket (schedule shape 1), one tile per
// workgroup, with order P's own LDS rows and registers.
#include <hip/hip_runtime.h>

#include "dg_mixed_tiled.h"

namespace dgp {

constexpr int MIX_BLOCK = 256;

int mixed_grid(int nchunks) { return nchunks < 2048 ? nchunks : 2048; }

template <int P, class LAW, int PU, bool PF>
__device__ __forceinline__ void mixed_tile(const MixedArgs& a, MixSmem<LAW, PU>& s, int tid, int i0, double& rsum) {
  using K = TiledMixed<P, LAW, PU, PF>;
  typename K::Regs g;
  K::phaseA(a, s, tid, i0, g);
  __syncthreads();
  for (int q0 = 0; q0 < K::NQ; q0 += MIX_QC) {
    K::phaseB(a, s, tid, q0);
    __syncthreads();
    K::phaseD1(a, s, tid, q0, g);
    __syncthreads();
  }
  K::phaseC1(a, s, tid, i0, g);
  __syncthreads();
  K::phaseC2(a, s, tid);
  __syncthreads();
  K::phaseD2(a, s, tid, i0, g);
  rsum += g.rsum;
}

template <class LAW>
__global__ __launch_bounds__(MIX_BLOCK) void tri_mixed_stage(MixedArgs a) {
  __shared__ MixSmem<LAW> sm;
  const int tid = threadIdx.x;
  double rsum = 0.0;
  TiledMixed<1, LAW>::stage_tables(a, sm, tid);  // the same tables for every order
  // the hardware hands workgroup b to XCD b % 8: numbering the workgroups of
  // one XCD consecutively gives each XCD a contiguous run of chunks (in mesh
  // order) at every pass, so the neighbour rows the tiles in flight gather are
  // in that XCD's L2
  const int nb = gridDim.x, full = (nb / 8) * 8, b = blockIdx.x;
  const int lw = b >= full ? b : (b % 8) * (nb / 8) + b / 8;
  for (int w = lw; w < a.nchunks; w += nb) {
    const int p = a.chunks[2 * w], i0 = a.chunks[2 * w + 1];
    switch (p) {
      case 1: mixed_tile<1, LAW, MAXP_TRI, false>(a, sm, tid, i0, rsum); break;
      case 2: mixed_tile<2, LAW, MAXP_TRI, false>(a, sm, tid, i0, rsum); break;
      case 3: mixed_tile<3, LAW, MAXP_TRI, false>(a, sm, tid, i0, rsum); break;
      default: mixed_tile<4, LAW, MAXP_TRI, false>(a, sm, tid, i0, rsum); break;
    }
    __syncthreads();  // the next tile overwrites the LDS
  }
  if (a.resid_part) {
    sm.red[tid] = rsum;
    __syncthreads();
    for (int s = MIX_BLOCK / 2; s > 0; s >>= 1) {
      if (tid < s) sm.red[tid] += sm.red[tid + s];
      __syncthreads();
    }
    if (tid == 0) a.resid_part[blockIdx.x] = sm.red[0];
  }
}

// One order bucket per launch: tile b takes the bucket's elements
// [bstart[P] + 64 b, + 64); the LDS and registers are those of order P alone.
// Residual partial sums at resid_part[part0 + b].
template <class LAW, int P>
__global__ __launch_bounds__(MIX_BLOCK) void tri_mixed_stage_p(MixedArgs a, int part0) {
  __shared__ MixSmem<LAW, P> sm;
  const int tid = threadIdx.x;
  double rsum = 0.0;
  TiledMixed<P, LAW, P, true>::stage_tables(a, sm, tid);
  __syncthreads();
  // consecutive tiles to one XCD (the hardware deals workgroups to the 8 XCDs
  // in turn), so the neighbour rows its tiles gather stay in that XCD's L2
  const int nb = gridDim.x, full = (nb / 8) * 8, bx = blockIdx.x;
  const int b = bx >= full ? bx : (bx % 8) * (nb / 8) + bx / 8;
  const int i0 = a.bstart[P] + b * MIX_TILE;
  mixed_tile<P, LAW, P, true>(a, sm, tid, i0, rsum);
  if (a.resid_part) {
    __syncthreads();
    sm.red[tid] = rsum;
    __syncthreads();
    for (int s = MIX_BLOCK / 2; s > 0; s >>= 1) {
      if (tid < s) sm.red[tid] += sm.red[tid + s];
      __syncthreads();
    }
    if (tid == 0) a.resid_part[part0 + b] = sm.red[0];
  }
}

template <class LAW>
__global__ __launch_bounds__(MIX_BLOCK) void tri_mixed_dt(MixedArgs a, double cfl, double* part) {
  const int i = blockIdx.x * MIX_BLOCK + threadIdx.x;
  const double dt = i < a.E ? dt_element<LAW>(a, i, cfl) : 1e300;
  __shared__ double red[MIX_BLOCK];
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = MIX_BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// one launch per order bucket: each order's limiter has its own registers
template <class LAW, int P>
__global__ __launch_bounds__(256) void tri_limit(LimiterArgs a) {
  const int i = a.bstart[P] + blockIdx.x * 256 + threadIdx.x;
  if (i < a.bstart[P + 1]) limit_element_p<LAW, P>(a, i);
}

// one launch per order bucket; a wave with any flagged lane sets the status once
template <int NF, bool EULER, int P>
__global__ __launch_bounds__(256) void tri_recover(RecoverArgs a) {
  const int i = a.l.bstart[P] + blockIdx.x * 256 + threadIdx.x;
  unsigned bits = 0;
  if (i < a.l.bstart[P + 1]) bits = recover_element_p<NF, EULER, P>(a, i);
  const unsigned b3 = (__any(bits & 1) ? 1u : 0u) | (__any(bits & 2) ? 2u : 0u) | (__any(bits & 4) ? 4u : 0u);
  if (b3 && (threadIdx.x & 63) == 0) atomicOr(a.status, b3);
}

template <int NF, bool EULER>
static int launch_recover_t(const RecoverArgs& a, hipStream_t s) {
  void (*kern[4])(RecoverArgs) = {tri_recover<NF, EULER, 1>, tri_recover<NF, EULER, 2>, tri_recover<NF, EULER, 3>,
                                  tri_recover<NF, EULER, 4>};
  for (int p = 1; p <= MAXP_TRI; p++) {
    const int n = a.l.bstart[p + 1] - a.l.bstart[p];
    if (n > 0) hipLaunchKernelGGL(kern[p - 1], dim3((n + 255) / 256), dim3(256), 0, s, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_recover(const RecoverArgs& a, int law, int nf, hipStream_t s) {
  if (law == LAW_EULER2D) return launch_recover_t<4, true>(a, s);
  if (nf == 1) return launch_recover_t<1, false>(a, s);
  if (nf == 3) return launch_recover_t<3, false>(a, s);
  return -1;
}

__global__ __launch_bounds__(256) void tri_sensor(SensorArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < a.E) sensor_element(a, i);
}

// order p (1..4) of element i and its bucket's first element and offset, by
// comparisons and select chains (no per-lane index into the argument block)
__device__ __forceinline__ void remap_bucket(const int* bs, const long* bo, int i, int& p, int& b0, int& b1, long& off) {
  p = 1 + (i >= bs[2]) + (i >= bs[3]) + (i >= bs[4]);
  b0 = p == 1 ? bs[1] : p == 2 ? bs[2] : p == 3 ? bs[3] : bs[4];
  b1 = p == 1 ? bs[2] : p == 2 ? bs[3] : p == 3 ? bs[4] : bs[5];
  off = p == 1 ? bo[1] : p == 2 ? bo[2] : p == 3 ? bo[3] : bo[4];
}

__global__ __launch_bounds__(256) void tri_remap(RemapArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.E) return;
  const int j = a.iold[i];
  int p1, s1, e1, p0, s0, e0;
  long o1, o0;
  remap_bucket(a.bs1, a.bo1, i, p1, s1, e1, o1);
  remap_bucket(a.bs0, a.bo0, j, p0, s0, e0, o0);
  const int np1 = (p1 + 1) * (p1 + 2) / 2, np0 = (p0 + 1) * (p0 + 2) / 2;
  const long n1 = e1 - s1, n0 = e0 - s0;
  const long d = o1 + (i - s1), r = o0 + (j - s0);
  for (int f = 0; f < a.NF; f++)
    for (int k = 0; k < np1; k++) a.dst[d + (long)(f * np1 + k) * n1] = k < np0 ? a.src[r + (long)(f * np0 + k) * n0] : 0.0;
  if (a.mg1) {
    for (int q = 0; q < 4; q++) a.mg1[(long)q * a.E + i] = a.mg0[(long)q * a.E + j];
    a.detj1[i] = a.detj0[j];
    a.size1[i] = a.size0[j];
  }
}

int launch_remap(const RemapArgs& a, hipStream_t s) {
  if (a.E <= 0) return -1;
  hipLaunchKernelGGL(tri_remap, dim3((a.E + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// grid: workgroups of the persistent launch (1 .. nchunks)
int launch_mixed_stage(const MixedArgs& a, int law, int ngrid, hipStream_t s) {
  if (a.nchunks <= 0 || ngrid <= 0 || ngrid > a.nchunks) return -1;
  dim3 grid(ngrid), block(MIX_BLOCK);
  if (law == LAW_SCALAR)
    hipLaunchKernelGGL(tri_mixed_stage<ScalarLawD>, grid, block, 0, s, a);
  else if (law == LAW_EULER2D && a.law.char_bc && a.law.flux == FLUX_EXACT)
    hipLaunchKernelGGL(tri_mixed_stage<EulerExactCharLawD>, grid, block, 0, s, a);
  else if (law == LAW_EULER2D && a.law.char_bc)
    hipLaunchKernelGGL(tri_mixed_stage<EulerCharLawD>, grid, block, 0, s, a);
  else if (law == LAW_EULER2D && a.law.flux == FLUX_EXACT)
    hipLaunchKernelGGL(tri_mixed_stage<EulerExactLawD>, grid, block, 0, s, a);
  else if (law == LAW_EULER2D)
    hipLaunchKernelGGL(tri_mixed_stage<EulerLawD>, grid, block, 0, s, a);
  else if (law == LAW_MAXWELL2D)
    hipLaunchKernelGGL(tri_mixed_stage<Maxwell2dLawD>, grid, block, 0, s, a);
  else if (law == LAW_BURGERS2D)
    hipLaunchKernelGGL(tri_mixed_stage<Burgers2DLawD>, grid, block, 0, s, a);
  else
    return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class LAW>
static int launch_mixed_bucket_t(const MixedArgs& a, int p, int part0, hipStream_t s) {
  void (*kern[4])(MixedArgs, int) = {tri_mixed_stage_p<LAW, 1>, tri_mixed_stage_p<LAW, 2>, tri_mixed_stage_p<LAW, 3>,
                                     tri_mixed_stage_p<LAW, 4>};
  const int n = mixed_bucket_tiles(a.bstart, p);
  if (n > 0) hipLaunchKernelGGL(kern[p - 1], dim3(n), dim3(MIX_BLOCK), 0, s, a, part0);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_mixed_bucket(const MixedArgs& a, int law, int p, int part0, hipStream_t s) {
  if (p < 1 || p > MAXP_TRI) return -1;
  if (law == LAW_SCALAR) return launch_mixed_bucket_t<ScalarLawD>(a, p, part0, s);
  if (law == LAW_EULER2D && a.law.char_bc && a.law.flux == FLUX_EXACT)
    return launch_mixed_bucket_t<EulerExactCharLawD>(a, p, part0, s);
  if (law == LAW_EULER2D && a.law.char_bc) return launch_mixed_bucket_t<EulerCharLawD>(a, p, part0, s);
  if (law == LAW_EULER2D && a.law.flux == FLUX_EXACT) return launch_mixed_bucket_t<EulerExactLawD>(a, p, part0, s);
  if (law == LAW_EULER2D) return launch_mixed_bucket_t<EulerLawD>(a, p, part0, s);
  if (law == LAW_MAXWELL2D) return launch_mixed_bucket_t<Maxwell2dLawD>(a, p, part0, s);
  if (law == LAW_BURGERS2D) return launch_mixed_bucket_t<Burgers2DLawD>(a, p, part0, s);
  return -1;
}

int mixed_dt_blocks(int E) { return (E + MIX_BLOCK - 1) / MIX_BLOCK; }

int launch_mixed_dt(const MixedArgs& a, int law, double cfl, double* part, hipStream_t s) {
  dim3 grid(mixed_dt_blocks(a.E)), block(MIX_BLOCK);
  if (law == LAW_SCALAR)
    hipLaunchKernelGGL(tri_mixed_dt<ScalarLawD>, grid, block, 0, s, a, cfl, part);
  else if (law == LAW_EULER2D)
    hipLaunchKernelGGL(tri_mixed_dt<EulerLawD>, grid, block, 0, s, a, cfl, part);
  else if (law == LAW_MAXWELL2D)
    hipLaunchKernelGGL(tri_mixed_dt<Maxwell2dLawD>, grid, block, 0, s, a, cfl, part);
  else if (law == LAW_BURGERS2D)
    hipLaunchKernelGGL(tri_mixed_dt<Burgers2DLawD>, grid, block, 0, s, a, cfl, part);
  else
    return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class LAW>
static void launch_limiter_law(const LimiterArgs& a, int p, hipStream_t s) {
  void (*kern[4])(LimiterArgs) = {tri_limit<LAW, 1>, tri_limit<LAW, 2>, tri_limit<LAW, 3>, tri_limit<LAW, 4>};
  const int n = a.bstart[p + 1] - a.bstart[p];
  if (n > 0) hipLaunchKernelGGL(kern[p - 1], dim3((n + 255) / 256), dim3(256), 0, s, a);
}

int launch_limiter_bucket(const LimiterArgs& a, int law, int p, hipStream_t s) {
  if (p < 1 || p > MAXP_TRI) return -1;
  if (law == LAW_SCALAR)
    launch_limiter_law<ScalarLawD>(a, p, s);
  else if (law == LAW_EULER2D && a.law.char_bc)
    launch_limiter_law<EulerCharLawD>(a, p, s);
  else if (law == LAW_EULER2D)
    launch_limiter_law<EulerLawD>(a, p, s);
  else
    return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_limiter(const LimiterArgs& a, int law, hipStream_t s) {
  for (int p = 1; p <= MAXP_TRI; p++)
    if (launch_limiter_bucket(a, law, p, s)) return -2;
  return 0;
}

int launch_sensor(const SensorArgs& a, hipStream_t s) {
  dim3 grid((a.E + 255) / 256), block(256);
  hipLaunchKernelGGL(tri_sensor, grid, block, 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace dgp
