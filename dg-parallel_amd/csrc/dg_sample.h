// Values of the modal state at the points of a regular lattice on every
// element's reference simplex: the interpolation step of the state export
// (DGCell::interpolate at the corners of the plot sub-simplices of
// DGAnalysis::exportGmsh, DGSensor.cc:337-525, and exportGmsh_Volume, :800-).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dgp {

constexpr int SAMPLE_MAX_NSUB = 8;

// Lattice (i, j[, k]) / nsub with i + j [+ k] <= nsub, in the order k (outer),
// j, i (inner); reference simplex (0,0),(1,0),(0,1) [,(0,0,1)].
inline int lattice_points(int dim, int nsub) {
  return dim == 2 ? (nsub + 1) * (nsub + 2) / 2 : (nsub + 1) * (nsub + 2) * (nsub + 3) / 6;
}

struct SampleArgs {
  int E, NF, NN;         // elements, fields, lattice points per element
  int bstart[6];         // order buckets of the internal numbering (bucket p = [bstart[p], bstart[p+1]))
  long boff[6];          // first double of bucket p in U ([row][element] within the bucket)
  int np[6];             // modes of order p
  int toff[6];           // basis of order p at the lattice points: tab[toff[p] + n * np[p] + k]
  const double* U;
  const double* tab;
  const int32_t* perm;   // internal element -> caller element
  double* out;           // [E][NF][NN] in caller element order
};

int launch_sample(const SampleArgs& a, hipStream_t s);

// Per-field totals int_K u_f = |J| U_f0 int phi_0 (the other modes integrate to
// zero) and energies int_K u_f^2 = |J| sum_k U_fk^2 (orthonormal modes), summed
// over the elements of each block of TOTALS_BLOCK; part[block][2 NF] (totals,
// then energies), added by the host in block order.
constexpr int TOTALS_BLOCK = 256;
struct TotalsArgs {
  int E, NF;
  int bstart[6];
  long boff[6];
  int np[6];
  double phi0_int;      // int over the reference element of phi_0
  const double* U;
  const double* detj;   // |J| per internal element
  double* part;
};
int launch_totals(const TotalsArgs& a, hipStream_t s);

}  // namespace dgp
