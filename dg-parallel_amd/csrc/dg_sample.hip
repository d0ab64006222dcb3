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
