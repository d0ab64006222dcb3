// Local time stepping: the FP64 tiled stage kernel over a list of tiles, each
// carrying its class's element range, stage flags, level and weight (one launch
// per stage of a step; the method is stated in dg_lts.h).
#include "dg_stage_tiled_kern.h"

namespace dgp {

template <int P, class LAW, int NW>
__global__ __launch_bounds__(NW * 64) void tri_stage_lts(StageArgs a) {
  using K = TiledStage<P, LAW, double, double, double, NW>;
  __shared__ typename K::Smem sm;
  const LtsTile tl = a.lts_tiles[xcd_tile(blockIdx.x, gridDim.x)];
  StageArgs b = a;
  b.e0 = tl.e0;
  b.e1 = tl.e1;
  b.stage = tl.flags;
  b.dt = a.dt / (double)(1 << tl.b);  // exact: a power of two
  b.lts_w = tl.w;
  b.resid_part = nullptr;
  tiled_body<K, false>(b, sm, 0);
}

template <class LAW, int NW>
static int launch_lts_nw(const StageArgs& a, int ntiles, hipStream_t s) {
  dim3 grid(ntiles), block(NW * 64);
  switch (a.P) {
    case 1: hipLaunchKernelGGL((tri_stage_lts<1, LAW, NW>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((tri_stage_lts<2, LAW, NW>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((tri_stage_lts<3, LAW, NW>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((tri_stage_lts<4, LAW, NW>), grid, block, 0, s, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// the 4-wave tile shape: launches of a few tiles take one tile's latency with
// either shape (about 17 us for P3 Euler, measured), the full ones are faster
template <class LAW>
static int launch_lts_law(const StageArgs& a, int ntiles, hipStream_t s) {
  return launch_lts_nw<LAW, 4>(a, ntiles, s);
}

// local time stepping: the span's increment added to y_m on the tiles'
// elements (one workgroup per tile: lane = element, waves stride the rows)
__global__ __launch_bounds__(256) void lts_commit_kernel(int E, int nrow, const LtsTile* __restrict__ tiles,
                                                         const double* __restrict__ acc, double* __restrict__ Un) {
  const LtsTile tl = tiles[blockIdx.x];
  const int e = tl.e0 + (threadIdx.x & 63);
  if (e >= tl.e1) return;
  for (int r = threadIdx.x >> 6; r < nrow; r += 4) Un[(long)r * E + e] += acc[(long)r * E + e];
}

int launch_lts_commit(int E, int nrow, const LtsTile* tiles, int ntiles, const double* acc, double* Un,
                      hipStream_t s) {
  if (ntiles <= 0) return 0;
  hipLaunchKernelGGL(lts_commit_kernel, dim3(ntiles), dim3(256), 0, s, E, nrow, tiles, acc, Un);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_tri_stage_lts(const StageArgs& a, int law, int ntiles, hipStream_t s) {
  if (ntiles <= 0) return 0;
  if (law == LAW_SCALAR && a.pvel) return launch_lts_law<ScalarFieldLawD>(a, ntiles, s);
  if (law == LAW_SCALAR) return launch_lts_law<ScalarLawD>(a, ntiles, s);
  if (law == LAW_EULER2D && a.law.char_bc)
    return a.law.flux == FLUX_EXACT ? launch_lts_law<EulerExactCharLawD>(a, ntiles, s)
                                    : launch_lts_law<EulerCharLawD>(a, ntiles, s);
  if (law == LAW_EULER2D && a.law.flux == FLUX_EXACT) return launch_lts_law<EulerExactLawD>(a, ntiles, s);
  if (law == LAW_EULER2D) return launch_lts_law<EulerLawD>(a, ntiles, s);
  if (law == LAW_MAXWELL2D) return launch_lts_law<Maxwell2dLawD>(a, ntiles, s);
  if (law == LAW_BURGERS2D) return launch_lts_law<Burgers2DLawD>(a, ntiles, s);
  return -1;
}

}  // namespace dgp
