// State export: every field at the lattice points of every element.
//
// One work-item per element.  For each field the element's Np coefficients are
// read once (coalesced: consecutive work-items are consecutive elements of a
// bucket's [row][element] arrays) and combined with the shared table of basis
// values at the lattice points.  Output goes to the caller's element order,
// NF * NN contiguous doubles per element.  The kernel is bound by HBM traffic
// (NF * (Np + NN) doubles per element), not by arithmetic.
#include "dg_sample.h"

namespace dgp {

template <int NP>
__device__ __forceinline__ void sample_element(const SampleArgs& a, int i, int p) {
  const int Ep = a.bstart[p + 1] - a.bstart[p];
  const double* U = a.U + a.boff[p] + (i - a.bstart[p]);
  const double* T = a.tab + a.toff[p];
  double* o = a.out + (size_t)a.perm[i] * a.NF * a.NN;
  for (int f = 0; f < a.NF; f++) {
    double c[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) c[k] = U[(size_t)(f * NP + k) * Ep];
    for (int n = 0; n < a.NN; n++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < NP; k++) s = fma(T[n * NP + k], c[k], s);
      o[f * a.NN + n] = s;
    }
  }
}

__global__ void __launch_bounds__(256) sample_lattice(SampleArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.E) return;
  // the bucket holding i: the largest p with bstart[p] <= i (empty buckets are
  // passed over because the next bucket starts at the same index)
  int p = 1;
#pragma unroll
  for (int q = 2; q <= 4; q++)
    if (i >= a.bstart[q]) p = q;
  switch (a.np[p]) {
    case 3: sample_element<3>(a, i, p); break;
    case 4: sample_element<4>(a, i, p); break;
    case 6: sample_element<6>(a, i, p); break;
    case 10: sample_element<10>(a, i, p); break;
    case 15: sample_element<15>(a, i, p); break;
    case 20: sample_element<20>(a, i, p); break;
    default: break;
  }
}

__global__ void __launch_bounds__(TOTALS_BLOCK) state_totals(TotalsArgs a) {
  __shared__ double red[TOTALS_BLOCK];
  const int i = blockIdx.x * TOTALS_BLOCK + threadIdx.x;
  double tot[6] = {0, 0, 0, 0, 0, 0}, en[6] = {0, 0, 0, 0, 0, 0};
  if (i < a.E) {
    int p = 1;
#pragma unroll
    for (int q = 2; q <= 4; q++)
      if (i >= a.bstart[q]) p = q;
    const int Ep = a.bstart[p + 1] - a.bstart[p], np = a.np[p];
    const double* U = a.U + a.boff[p] + (i - a.bstart[p]);
    const double dj = a.detj[i];
    for (int f = 0; f < a.NF && f < 6; f++) {
      double s2 = 0.0;
      for (int k = 0; k < np; k++) {
        const double v = U[(size_t)(f * np + k) * Ep];
        s2 = fma(v, v, s2);
      }
      tot[f] = dj * a.phi0_int * U[(size_t)(f * np) * Ep];
      en[f] = dj * s2;
    }
  }
  // one fixed-order tree per quantity
  for (int j = 0; j < 2 * a.NF; j++) {
    const int f = j < a.NF ? j : j - a.NF;
    double v = 0.0;
#pragma unroll
    for (int g = 0; g < 6; g++)
      if (g == f) v = j < a.NF ? tot[g] : en[g];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = TOTALS_BLOCK / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.part[(size_t)blockIdx.x * 2 * a.NF + j] = red[0];
    __syncthreads();
  }
}

int launch_totals(const TotalsArgs& a, hipStream_t s) {
  if (a.E <= 0 || a.NF < 1 || a.NF > 6) return -1;
  hipLaunchKernelGGL(state_totals, dim3((a.E + TOTALS_BLOCK - 1) / TOTALS_BLOCK), dim3(TOTALS_BLOCK), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_sample(const SampleArgs& a, hipStream_t s) {
  if (a.E <= 0) return 0;
  for (int p = 1; p <= 4; p++) {
    if (a.bstart[p + 1] == a.bstart[p]) continue;
    const int n = a.np[p];
    if (n != 3 && n != 4 && n != 6 && n != 10 && n != 15 && n != 20) return -1;
  }
  hipLaunchKernelGGL(sample_lattice, dim3((a.E + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace dgp
