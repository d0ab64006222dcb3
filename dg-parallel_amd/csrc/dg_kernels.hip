// Hand-written gfx950 kernels for the explicit DG residual and Runge-Kutta
// stage on triangles.
//
// One fused kernel per RK stage (tri_stage_kernel): each work-item owns one
// element, keeps its modal coefficients and residual in registers, and
//   1. volume term (DGCell.cc:184-249): at every volume point interpolates the
//      state, evaluates the physical flux, maps it with |J| J^-1 (one 2x2 per
//      element) and contracts with the reference basis gradients;
//   2. face term (DGCell.cc:639-732): for each of its three faces evaluates its
//      own trace and the neighbour's trace (neighbour coefficients read straight
//      from the stage-input array, which no kernel writes during the stage),
//      the numerical flux, and lifts it with its own trace basis.  Each face
//      flux is evaluated by both neighbours: no face buffer, no atomics, and
//      the result does not depend on scheduling;
//   3. stage update (TimeIntegrator.cc:81-275, 279-438, 838-990): applies the
//      inverse mass 1/|J| and the RK stage formula, reading/writing the FP64
//      step-start state and accumulator and writing the next stage input.
// Tables (basis values at quadrature points) are addressed with compile-time
// offsets from a kernel-argument pointer, so they are wave-uniform scalar loads.
// Elements are ordered on the host so that a wavefront shares the neighbour's
// local-face index on each face; the switch below then runs one path per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

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

template <int P, class LAW, typename CT, typename IT, int LB, bool SAME>
__device__ __forceinline__ void nb_traces(const IT* __restrict__ Uin, int E, int nb, const CT* __restrict__ tb,
                                          CT (&uR)[nqf_tri(P)][LAW::NF]) {
  constexpr int NP = np_tri(P), NQF = nqf_tri(P), NF = LAW::NF;
  constexpr int OFF = TriBlock::off_phif(P);
#pragma unroll
  for (int f = 0; f < NF; f++) {
    CT v[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) v[k] = CT(Uin[(long)(f * NP + k) * E + nb]);
#pragma unroll
    for (int q = 0; q < NQF; q++) {
      constexpr int dummy = 0;
      (void)dummy;
      const int qq = SAME ? q : NQF - 1 - q;
      CT s = CT(0);
#pragma unroll
      for (int k = 0; k < NP; k++) s += tb[OFF + (LB * NQF + qq) * NP + k] * v[k];
      uR[q][f] = s;
    }
  }
}

template <int P, class LAW, typename CT, typename IT>
__device__ __forceinline__ void nb_traces_any(const IT* __restrict__ Uin, int E, int link, const CT* __restrict__ tb,
                                              CT (&uR)[nqf_tri(P)][LAW::NF]) {
  int nb = link >> 3;
  int lb = link & 3;
  bool same = (link >> 2) & 1;
  if (!same) {
    if (lb == 0)
      nb_traces<P, LAW, CT, IT, 0, false>(Uin, E, nb, tb, uR);
    else if (lb == 1)
      nb_traces<P, LAW, CT, IT, 1, false>(Uin, E, nb, tb, uR);
    else
      nb_traces<P, LAW, CT, IT, 2, false>(Uin, E, nb, tb, uR);
  } else {
    if (lb == 0)
      nb_traces<P, LAW, CT, IT, 0, true>(Uin, E, nb, tb, uR);
    else if (lb == 1)
      nb_traces<P, LAW, CT, IT, 1, true>(Uin, E, nb, tb, uR);
    else
      nb_traces<P, LAW, CT, IT, 2, true>(Uin, E, nb, tb, uR);
  }
}

template <int P, class LAW, typename CT, typename IT, typename OT>
__global__ __launch_bounds__(BLOCK) void tri_stage_kernel(StageArgs a) {
  constexpr int NP = np_tri(P), NQ = nqv_tri(P), NQF = nqf_tri(P), NF = LAW::NF;
  const int E = a.E;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int e = blk * BLOCK + threadIdx.x;
  double rsum = 0.0;
  if (e < E) {
    const CT* __restrict__ tb = sizeof(CT) == 8 ? (const CT*)a.tab : (const CT*)a.tabf;
    const IT* __restrict__ Uin = (const IT*)a.Uin;
    const LawParams& law = a.law;

    CT U[NF][NP];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int k = 0; k < NP; k++) U[f][k] = CT(Uin[(long)(f * NP + k) * E + e]);
    CT mg[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
      for (int d = 0; d < 2; d++) mg[r][d] = CT(a.el.mg[(long)(r * 2 + d) * E + e]);

    CT R[NF][NP];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int k = 0; k < NP; k++) R[f][k] = CT(0);

    // ---- volume term -------------------------------------------------------
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      CT u[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        CT s = CT(0);
#pragma unroll
        for (int k = 0; k < NP; k++) s += tb[TriBlock::off_phi(P) + q * NP + k] * U[f][k];
        u[f] = s;
      }
      CT F[NF][2];
      LAW::flux(law, u, F);
      const CT w = tb[TriBlock::off_wv(P) + q];
      CT G[NF][2];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        G[f][0] = w * (mg[0][0] * F[f][0] + mg[0][1] * F[f][1]);
        G[f][1] = w * (mg[1][0] * F[f][0] + mg[1][1] * F[f][1]);
      }
#pragma unroll
      for (int k = 0; k < NP; k++) {
        const CT g0 = tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 0];
        const CT g1 = tb[TriBlock::off_dphi(P) + (q * NP + k) * 2 + 1];
#pragma unroll
        for (int f = 0; f < NF; f++) R[f][k] += g0 * G[f][0] + g1 * G[f][1];
      }
    }

    // ---- face terms ----------------------------------------------------------
#pragma unroll
    for (int l = 0; l < 3; l++) {
      const int link = a.el.link[(long)l * E + e];
      // outward normal times edge length: |J| J^-T nhat |fhat| (nhat |fhat| of the
      // reference edges: (0,-1), (1,1), (-1,0))
      CT nl0, nl1;
      if (l == 0) {
        nl0 = -mg[1][0];
        nl1 = -mg[1][1];
      } else if (l == 1) {
        nl0 = mg[0][0] + mg[1][0];
        nl1 = mg[0][1] + mg[1][1];
      } else {
        nl0 = -mg[0][0];
        nl1 = -mg[0][1];
      }
      const CT len = dg_sqrt(nl0 * nl0 + nl1 * nl1);
      const CT n[2] = {nl0 / len, nl1 / len};
      CT uR[NQF][NF];
      if (link >= 0) nb_traces_any<P, LAW, CT, IT>(Uin, E, link, tb, uR);
#pragma unroll
      for (int q = 0; q < NQF; q++) {
        CT uL[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          CT s = CT(0);
#pragma unroll
          for (int k = 0; k < NP; k++) s += tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k] * U[f][k];
          uL[f] = s;
        }
        CT Fh[NF];
        face_flux<LAW>(law, link, n, uL, uR[q], Fh);
        const CT sw = len * tb[TriBlock::off_wf(P) + q];  // |J_face| w = (len/2) w
#pragma unroll
        for (int f = 0; f < NF; f++) {
          const CT fs = Fh[f] * sw;
#pragma unroll
          for (int k = 0; k < NP; k++) R[f][k] -= fs * tb[TriBlock::off_phif(P) + (l * NQF + q) * NP + k];
        }
      }
    }

    // ---- stage update ------------------------------------------------------
    if (a.scheme == MODE_RESIDUAL) {
#pragma unroll
      for (int f = 0; f < NF; f++)
#pragma unroll
        for (int k = 0; k < NP; k++) a.R[(long)(f * NP + k) * E + e] = double(R[f][k]);
    } else {
      const double dtj = a.dt / a.el.detj[e];
      OT* __restrict__ Uout = (OT*)a.Uout;
      const double* __restrict__ Un = a.Un;
      double* __restrict__ acc = a.acc;
      const int sc = a.scheme * 4 + a.stage;
#pragma unroll
      for (int f = 0; f < NF; f++)
#pragma unroll
        for (int k = 0; k < NP; k++) {
          const long i = (long)(f * NP + k) * E + e;
          const double v = double(R[f][k]) * dtj;
          double o;
          switch (sc) {
            case 3 * 4 + 0: {  // RK4 stage 1 (TimeIntegrator.cc:101-137)
              acc[i] = v;
              o = Un[i] + 0.5 * v;
              break;
            }
            case 3 * 4 + 1: {  // RK4 stage 2 (:149-181)
              acc[i] += 2. * v;
              o = Un[i] + 0.5 * v;
              break;
            }
            case 3 * 4 + 2: {  // RK4 stage 3 (:192-222)
              acc[i] += 2. * v;
              o = Un[i] + v;
              break;
            }
            case 3 * 4 + 3: {  // RK4 stage 4 (:235-270)
              const double s = acc[i] + v;
              o = Un[i] + s / 6.0;
              rsum += s * s / 36.0;
              break;
            }
            case 2 * 4 + 0:  // SSP-RK3 stage 1 (:850-880)
              o = Un[i] + v;
              break;
            case 2 * 4 + 1:  // SSP-RK3 stage 2 (:889-925)
              o = 0.25 * (double(Uin[i]) + v + 3. * Un[i]);
              break;
            case 2 * 4 + 2: {  // SSP-RK3 stage 3 (:934-987)
              o = (2.0 * (double(Uin[i]) + v) + Un[i]) / 3.0;
              const double d = o - Un[i];
              rsum += d * d;
              break;
            }
            case 1 * 4 + 0:  // SSP-RK2 stage 1 (parallel.cc:78-131)
              o = Un[i] + v;
              break;
            case 1 * 4 + 1: {  // SSP-RK2 stage 2 (parallel.cc:133-186)
              o = (double(Uin[i]) + v + Un[i]) * 0.5;
              const double d = o - Un[i];
              rsum += d * d;
              break;
            }
            default: {  // forward Euler (TimeIntegrator.cc:993-1048)
              o = Un[i] + v;
              const double d = Un[i] - o;
              rsum += d * d;
              break;
            }
          }
          Uout[i] = OT(o);
        }
    }
  }
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

int launch_copy_probe(const void* src, void* dst, long n_doubles, hipStream_t s) {
  long n4 = n_doubles / 4;
  hipLaunchKernelGGL(copy_probe_kernel, dim3(4096), dim3(256), 0, s, (const double4*)src, (double4*)dst, n4);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace dgp
