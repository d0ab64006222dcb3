// Launch interface between the host context (dg_capi.cc) and the HIP kernels
// (dg_kernels.hip).  Plain structs of device pointers; no torch types.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dg_laws.h"
#include "dg_lts.h"

namespace dgp {

// Element table of one order bucket of triangles, structure-of-arrays over
// the bucket's E elements (element index fastest => coalesced per wavefront).
struct TriElemTable {
  const double* mg;     // [4][E]  mg[r*2+d] = |J| d xi_r / d x_d
  const double* detj;   // [E]     |J| (the inverse mass of the orthonormal basis is 1/|J|)
  const double* size;   // [E]     2 x inradius (DGCell.cc:1209-1229)
  const int32_t* link;  // [3][E]  neighbour across local face l:
                        //         >= 0: (nb << 3) | (same_orientation << 2) | nb_local_face
                        //         <  0: -(boundary code)
};

// stage modes
enum { MODE_RESIDUAL = -1 };
// the stage of local time stepping (dg_lts.h): stage holds the LTS_* flags, dt
// the level's substep h_b and lts_w the weight of h_b k in the span increment
enum { SCHEME_LTS = 4 };

struct StageArgs {
  int E;
  int P;
  const void* Uin;       // stage input, SoA [NF*NP][E] (double or float)
  const double* Un;      // step-start state, SoA, FP64
  double* acc;           // RK4 accumulator, FP64
  void* Uout;            // stage output (double or float)
  double* R;             // MODE_RESIDUAL output, SoA FP64
  const double* tab;     // table block of order P (FP64)
  const float* tabf;     // same block in FP32
  TriElemTable el;
  LawParams law;
  double dt;
  int scheme;            // DG_RK_* or MODE_RESIDUAL
  int stage;
  double* resid_part;    // optional per-block partial sums of the step residual norm
  // the scalar law's velocity at the quadrature points (a velocity field), FP64,
  // [NQV + 3 NQF][2][E]: volume points, then face l point q at NQV + l NQF + q
  const double* pvel;
  // the tiled kernels update internal elements [e0, e1) only (e1 = 0: all E);
  // every element's rows keep the stride E
  int e0, e1;
  double lts_w;
  // local time stepping: the launch's tiles (launch_tri_stage_lts), each with
  // its own range, flags (stage), level (dt / 2^b) and weight (lts_w)
  const LtsTile* lts_tiles;
  // FP32 form, a stage with FP64 output (the step's last): the output rounded
  // to FP32 as well, the next step's first-stage input (null: not written)
  float* Uout32;
};

// variant: 0 = FP64 throughout; 1 = FP32 arithmetic, FP64 input, FP32 output;
// 2 = FP32 in/out; 3 = FP32 input, FP64 output
int launch_tri_stage(const StageArgs& a, int law, int variant, hipStream_t s);
int tri_stage_blocks(int E);
// tiled kernel (64 elements per workgroup, LDS-staged); same arguments
int launch_tri_stage_tiled(const StageArgs& a, int law, int variant, hipStream_t s);
int tri_tiled_blocks(int E);

int launch_copy_probe(const void* src, void* dst, long n_doubles, hipStream_t s);
int launch_round_f32(const double* in, float* out, long n, hipStream_t s);

struct AdjArgs;
int launch_tri_vjp(const AdjArgs& a, int law, int nd, hipStream_t s);
int launch_tri_jvp(const AdjArgs& a, int law, hipStream_t s);
int launch_tri_adj_scale(int E, int nrow, double c, double dt, const double* detj, const double* in, double* out,
                         hipStream_t s);

struct DtArgs {
  int E, P;
  const double* U;       // SoA FP64
  const double* vtx;     // [4][NP]: phi at the centroid slot 0 (phi_0 only) then 3 vertices
  TriElemTable el;
  LawParams law;
  double cfl;
  double* part;          // per-block minima
  const double* pvel;    // the velocity table of StageArgs (or null)
  const double* dtscale; // per-element factor of dt (local time stepping: 2^level), or null
};
int launch_tri_dt(const DtArgs& a, int law, hipStream_t s);
int tri_dt_blocks(int E);

// The allocation-free time step: per-workgroup minima go to a file-scope
// device array of NOALLOC_SCRATCH_DOUBLES doubles (one copy per device, kept
// until the process ends) instead of a context buffer.  noalloc_scratch_init
// looks up that array's device address once per device (call with the device
// current); noalloc_scratch returns it (nullptr before the init).
constexpr int NOALLOC_SCRATCH_DOUBLES = 1 << 16;
int noalloc_scratch_init(int device);
double* noalloc_scratch(int device);
// out[0] = min(part[0 .. n)), one workgroup; out[0] = v (one work-item)
int launch_min_reduce(const double* part, int n, double* out, hipStream_t s);
// *status |= 4 when any of U[0 .. n) is NaN or Inf (one atomic per wave that meets one)
int launch_nonfinite_scan(const double* U, long n, unsigned* status, hipStream_t s);
int launch_fill_one(double* out, double v, hipStream_t s);
// Meshes with hanging nodes: after each stage kernel, one workgroup per element
// with faces of half an edge integrates them (both sides evaluated at the
// half's Gauss points, the interior two-point flux, FP64) and adds the
// contribution to the residual (MODE_RESIDUAL) or, through the stage update's
// coefficient of dt R / |J|, to the stage output and the RK4 accumulator.
constexpr int HANG_LANES = 64;   // work-items per element of the list
constexpr int HANG_MAXREC = 6;   // face halves per element: two on each face
struct HangArgs {
  int E, P, n;            // n: elements with faces of half an edge
  const int32_t* elem;    // [n] internal element
  const int32_t* rstart;  // [n + 1] into rec
  // per face of half an edge and side: own local face, the other side's
  // internal element, its local face, flags: bit 0 the own side is the coarse
  // cell, bit 1 the child runs against the coarse face's direction, bit 2 the half
  const int32_t* rec;     // [nrec][4]
  const double* Uin;      // stage input (FP64)
  double* Uout;
  double* acc;
  double* R;
  const double* tab;      // the order-P block (FP64)
  const double* phih;     // [3][2][NQF][NP]: the basis at the Gauss points of each face's halves
  TriElemTable el;
  LawParams law;
  double dt;
  int scheme, stage;
};
int launch_hang(const HangArgs& a, int law, hipStream_t s);

// local time stepping: Un += acc on rows [0, nrow) of the tiles' elements
int launch_lts_commit(int E, int nrow, const LtsTile* tiles, int ntiles, const double* acc, double* Un,
                      hipStream_t s);
// local time stepping: the FP64 tiled stage kernel (4- or 8-wave tiles) over a tile list
// (a.lts_tiles, ntiles workgroups), step size a.dt
int launch_tri_stage_lts(const StageArgs& a, int law, int ntiles, hipStream_t s);

}  // namespace dgp
