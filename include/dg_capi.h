/*
 * dg_capi.h -- C ABI of the MI355X (gfx950) DG residual / Runge-Kutta library.
 *
 * The drop-in boundary for DG-Parallel's explicit DG hot path: everything the
 * reference reaches through DGAnalysis (DGAnalysis.h:110-135),
 * TimeIntegrator::advanceInTime (TimeIntegrator.h:28) and the ConservationLaw
 * plug-in surface (ConservationLaw.h:12-60) is served here by hand-written HIP
 * kernels on one GPU.  Plain C types only: pointers, sizes, enums.
 *
 * Conventions
 *  - Caller owns host buffers; the context owns all device memory and one HIP
 *    stream.  Calls on one context are serialised on that stream; a context is
 *    used from one host thread at a time.
 *  - No C++ exception crosses the ABI and nothing calls exit(): every entry
 *    point returns a dg_status; dg_last_error() gives the message.
 *  - State layout at the boundary: element-major, field, mode -- U[e][f][k]
 *    with Np(e) modes, elements in the caller's order; the reference's cell
 *    coefficient block `i*FS+j` (DGCell.h:31) concatenated over cells.
 */
#ifndef DG_CAPI_H
#define DG_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dg_ctx dg_ctx;

typedef enum {
  DG_OK = 0,
  DG_ERR_ARG = 1,          /* bad argument / malformed mesh */
  DG_ERR_OOM = 2,          /* device allocation failed (nothing leaked) */
  DG_ERR_HIP = 3,          /* HIP runtime error */
  DG_ERR_NONFINITE = 4,    /* NaN / Inf in a state or residual */
  DG_ERR_NONPHYSICAL = 5,  /* negative density/pressure at a quadrature point */
  DG_ERR_UNSUPPORTED = 6   /* law / order / element combination not built */
} dg_status;

/* DG_LAW_MAXWELL2D: Maxwell2d, the TE fields (Hx, Hy, Ez) with its own upwind
 * interface state (MaxwellLaw.cc:21-200; flux DG_FLUX_UPWIND; walls mirror H);
 * DG_LAW_BURGERS2D: Burgers2D, F = (u^2/2, u^2/2) with its central flux
 * (Burgers2D.cc:22-92; flux DG_FLUX_CENTRAL; far field upwind on boundaries) */
enum { DG_LAW_SCALAR = 0, DG_LAW_EULER2D = 1, DG_LAW_MAXWELL3D = 2, DG_LAW_MAXWELL2D = 3, DG_LAW_BURGERS2D = 4 };
/* DG_FLUX_ROE is Euler2dRoe::riemannSolver (EulerLaw.cc:913-1010); DG_FLUX_ROE_EF
 * adds Harten's entropy fix (delta = 0.2 c), which the reference lacks;
 * DG_FLUX_EXACT is Euler2d's own riemannSolver, the exact Godunov flux
 * (EulerLaw.cc:602-887: Newton for p*, at most 12 iterations to 1e-14) */
enum { DG_FLUX_UPWIND = 0, DG_FLUX_LF = 1, DG_FLUX_ROE = 2, DG_FLUX_CENTRAL = 3, DG_FLUX_ROE_EF = 4,
       DG_FLUX_EXACT = 5 };
/* DG_RK_TVD2_LTS: RungeKuttaTVD2adaptive, SSP-RK2 with local time steps per
 * size level (TimeIntegrator.cc:442-833); needs dg_lts_setup */
enum { DG_RK_FE = 0, DG_RK_TVD2 = 1, DG_RK_TVD3 = 2, DG_RK4 = 3, DG_RK_TVD2_LTS = 4 };
enum { DG_PREC_FP64 = 0, DG_PREC_FP32 = 1 };
/* Boundary tags (the reference's classification ids, EulerLaw.cc:182-296):
 * 40000 reflective wall; 20000 the reference's curved wall (EulerLaw.cc:109-179,
 * its ghost state reflects the velocity about the exact surface normal N; on the
 * straight-sided faces built here N is the face normal and the ghost is the
 * 40000 one); 30000 characteristic far field (EulerLaw.cc:185-278: the
 * free stream or the interior state at supersonic normal speed, else the state
 * of the Riemann invariants J+ from outside and J- from inside; Euler, FP64 and
 * FP32 forms on the 4- and 8-wave tiles);
 * any other tag on a face without neighbour: Riemann flux against the law's
 * far-field state (the reference's default branch, EulerLaw.cc:293-295).  The
 * reference evaluates its far-field FieldEvaluator at the boundary point; tag
 * DG_TAG_FARFIELD2 selects a second constant state (Euler: farfield[4..8)), so
 * two boundaries can hold different states (the two ends of a shock tube). */
enum { DG_TAG_FARFIELD = 1, DG_TAG_FARFIELD2 = 2, DG_TAG_CURVED_WALL = 20000, DG_TAG_CHARACTERISTIC = 30000,
       DG_TAG_WALL = 40000 };

/* Mesh handed over by the caller (replaces mDGMesh + mImportExport for the path:
 * mMesh.h:52-139).  Faces are matched by their sorted topological vertex ids;
 * elements that share a vertex id through a periodic image are neighbours
 * (translation periodicity). */
typedef struct {
  int dim;                /* 2: triangles, 3: tetrahedra */
  int n_elems;
  const int32_t* vids;    /* [E][dim+1] topological vertex ids */
  const double* xyz;      /* [E][dim+1][dim] corner coordinates of each element */
  const int32_t* bc;      /* [E][dim+1] tag of faces with no neighbour, NULL: far field.
                             Local face l of a triangle is edge (v_l, v_{l+1}) (Face.cc:114-148) */
  const int32_t* order;   /* [E] polynomial order per element (1..4), NULL: options.order.
                             Mixed orders (DGCell::adaptOrder, DGCell.cc:1542-1602) run
                             the FP64 form in one persistent launch over order buckets;
                             the state layout is then [e][f][k] with Np(p_e) modes */
} dg_mesh_desc;

/* Law selection (replaces the ConservationLaw subclass instance; the host C++
 * classes in dg-parallel_amd/host map themselves onto this). */
typedef struct {
  int law;                /* DG_LAW_* */
  int flux;               /* DG_FLUX_*: Scalar upwind; Euler LF or Roe; Maxwell upwind or central */
  double gamma;           /* Euler */
  double vel[3];          /* ScalarLaw constant velocity field */
  double farfield[8];     /* ghost state of far-field boundary faces */
} dg_law_desc;

typedef struct {
  int order;              /* polynomial order when mesh.order is NULL */
  int precision;          /* DG_PREC_FP64: plain form; DG_PREC_FP32: FP32 residual, FP64 stage accumulation */
  int use_graph;          /* capture dg_advance's steps in hipGraphs and replay them (calls of 64 steps or
                             more: eight steps per graph, one-step graphs for the remainder) */
  int device;             /* HIP device ordinal */
  int reserved[8];        /* reserved[0]: stage kernel, 0 tiled (default), 1 one element per work-item;
                             reserved[1]: adjoint flux tangent directions per pass, 0 (default), 1, 2, 4 */
} dg_options;

/* DGAnalysis::DGAnalysis(mesh, law, limiter, order) (DGAnalysis.cc:133-302):
 * builds element/face tables on the host in FP64 and uploads them. */
dg_status dg_create(const dg_mesh_desc* mesh, const dg_law_desc* law, const dg_options* opt, dg_ctx** out);
/* DGAnalysis::~DGAnalysis (DGAnalysis.cc:304-307) */
dg_status dg_destroy(dg_ctx* ctx);
const char* dg_last_error(const dg_ctx* ctx);
/* sum_e Nf*Np(e): the length of every state / residual buffer */
size_t dg_ndof(const dg_ctx* ctx);
int dg_nfaces(const dg_ctx* ctx);
/* faces as (elemL, localL, elemR, localR) in caller element numbering, elemL <
 * elemR; boundary faces carry elemR = -1.  Order: sorted by sorted vertex ids. */
dg_status dg_faces(const dg_ctx* ctx, int32_t* out4, size_t n);

/* A FieldEvaluator (FieldEvaluator.h:12, eval(const mPoint&, double t, double* val)):
 * val receives the field's components at the physical point x (x[2] = 0 in 2D). */
typedef void (*dg_field_fn)(const double* x, double t, double* val, void* user);
/* ScalarLaw's velocity FieldEvaluator (ScalarLaw(farField, velocityField, exact),
 * ScalarLaw.cc:20-97: Fi, riemannSolver, boundaryFlux and maximumEigenValue
 * evaluate velocityField at the point they work on).  fn is called once per
 * volume and face quadrature point of every element, at time 0 as the
 * reference's Fi and riemannSolver call it, and the velocities are kept on the
 * device; the dt rule takes |vx| + |vy| at each element's first volume point
 * (DGCell::computeMAXEV, DGCell.cc:1165-1186).  Each element evaluates a face
 * point at its own position, so on a periodic face the two images must give the
 * same normal velocity (a periodic field, or a rotation about the centre of an
 * origin-centred square, as the reference's ScalarPb, Lilia.cc:66-108); the
 * reference takes one side's point.  Scalar law on triangles, uniform order, the
 * tiled kernel; derivatives, the adjoint and order changes refuse such a
 * context (DG_ERR_UNSUPPORTED).  Replaces the constant dg_law_desc.vel. */
dg_status dg_set_velocity_field(dg_ctx* ctx, dg_field_fn fn, void* user);

/* DGCell::theFieldsCoefficients write/read for all cells (DGCell.h:22-40) */
dg_status dg_set_state(dg_ctx* ctx, const double* U, size_t n);
dg_status dg_get_state(dg_ctx* ctx, double* U, size_t n);

/* One residual evaluation R = int F.grad(phi) - oint Fhat phi for all cells:
 * TimeIntegrator::assembleVolume + assembleBoundary (TimeIntegrator.cc:1332-1412)
 * = DGCell::computeVolumeContribution (DGCell.cc:184-249) over cells plus
 * DGBoundaryCell::computeBoundaryContributions (DGCell.cc:639-732) over faces.
 * R_out: host buffer of dg_ndof() doubles (NULL: keep on device). */
dg_status dg_residual(dg_ctx* ctx, double t, double* R_out);

/* DGCell::adaptTimeStep min over cells (DGCell.cc:1155-1262, DGAnalysis.cc:510-540):
 * dt = min_e CFL*size_e / ((2p_e+1) lambda_e), without the TEND snapping. */
dg_status dg_compute_dt(dg_ctx* ctx, double cfl, double* dt_out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same global
 * minimum as dg_compute_dt, written to dt_dev (a device pointer to one double)
 * on the context's stream, with no allocation, no host copy and no
 * synchronisation (the caller synchronises before reading it).  The
 * per-workgroup minima go to a fixed device array of 65 536 doubles (512 KiB,
 * one per device, held until the process ends), so triangle meshes are limited
 * to 16 777 216 elements (DG_ERR_ARG beyond); tetrahedra need no scratch.
 * Calls that overlap on one device (from several host threads or streams) race
 * on that array. */
dg_status dg_compute_dt_noalloc(dg_ctx* ctx, double cfl, double* dt_dev);

/* nsteps calls of TimeIntegrator::advanceInTime(t, dt) (TimeIntegrator.h:28) with
 * the scheme chosen by rk_scheme (ForwardEuler, RungeKuttaTVD2, RungeKuttaTVD3,
 * RungeKutta4: TimeIntegrator.cc:81-275,279-438,838-990,993-1048).  resid_out
 * (optional) receives the last step's residual norm as the reference returns it. */
dg_status dg_advance(dg_ctx* ctx, int rk_scheme, double t, double dt, int nsteps, double* resid_out);

/* ---- local time stepping (triangles, FP64 form, uniform order) ------------
 * DGAnalysis::sortCellsBySize (DGAnalysis.cc:749-890) with the integrator it
 * selects, RungeKuttaTVD2adaptive (DGAnalysis.cc:232-234; TimeIntegrator.cc:
 * 442-833).  Cells are binned by size (DGCell::getSize, 2 x inradius): level b
 * holds the cells with size <= 1.0001 smax / 2^b (the largest b below
 * max_levels, 1..6) and takes 2^b SSP-RK2 substeps per step; face neighbours
 * are kept within one level of each other by moving the coarser one a level
 * down (the reference stops with "Too big a difference between elements").
 * The reference's range assembly is empty (TimeIntegrator.cc:1438-1487), so its
 * adaptive step advances nothing; this build's multirate step is a new design:
 * a partitioned Runge-Kutta method of 2^levels stages whose levels share the
 * stage weights, so it is conservative and second order, and on one level it is
 * SSP-RK2.  Elements away from finer ones skip the stages whose inputs repeat.
 * After this call (which re-lays out the elements by level, keeping the state),
 * dg_compute_dt returns the step of the coarsest level, min over cells of their
 * CFL step x 2^level (the reference's dt x 2^(bins-1), DGAnalysis.cc:562-589),
 * and dg_advance / dg_time_advance take DG_RK_TVD2_LTS steps of that size.
 * levels_out (caller order, may be NULL) receives each cell's level; the return
 * is DG_OK, or DG_ERR_UNSUPPORTED for tetrahedra, mixed orders, the FP32 form,
 * the one-element kernel or a context with a velocity field (set the field after
 * this call).  The limiter and the recovery pass are not applied by
 * DG_RK_TVD2_LTS steps (DG_ERR_UNSUPPORTED when enabled).  max_levels = 1 puts
 * every cell on level 0. */
dg_status dg_lts_setup(dg_ctx* ctx, int max_levels, int32_t* levels_out);

/* ---- h-adaptivity (triangles, FP64 form, one order) -----------------------
 * DGAnalysis::adaptH's mesh operations (DGPref.cc:78-192): cells marked > 0 are
 * split into four through their edge midpoints (mDGMesh::splitTriangle,
 * mDGMesh.cc:234-337) while below max_level refinements (LEVEL_OF_REFINEMENT),
 * and families of four children all marked < 0 are merged back into their
 * parent (mDGMesh::unsplit, mDGMesh.cc:463-552); the state moves by exact L2
 * projection (DGAnalysis::splitCell / unsplitCell, DGPref.cc:632-671).  The
 * reference chooses the marks from its error estimate (modifyP, DGPref.cc:
 * 27-76); here the caller passes them (one per current cell, caller order).
 * Neighbours of split cells stay unsplit (hanging nodes: each of their split
 * faces is integrated over its two halves, as the reference's boundary cells of
 * leaf edges do, DGCell.cc:423-555); a coarser neighbour of a split cell is
 * split too, and a merge that would put a cell beside one two levels finer is
 * skipped, so every face meets at most two halves.  The context is rebuilt on
 * the new mesh: split cells give their place to their four children
 * (corners (P0, M01, M20), (M01, P1, M12), (M20, M12, P2), (M12, M20, M01)),
 * a merged family's parent takes its first child's place, the state, orders
 * and dg_get_state's layout follow the new cells; a new midpoint takes the next
 * free vertex id (the same id for the edge's other side and for later splits).
 * n_elems_out (may be NULL) receives the new cell count.  Meshes with hanging
 * nodes run the scalar, Euler, Maxwell2d and Burgers2D laws with every RK
 * scheme; the limiter, recovery, local time stepping, order changes, velocity
 * fields and residual derivatives refuse them (DG_ERR_UNSUPPORTED). */
dg_status dg_adapt_h(dg_ctx* ctx, const int8_t* marks, int max_level, int32_t* n_elems_out);
/* the current mesh (after dg_adapt_h: caller order): the cell count, corner ids
 * [E][3], coordinates [E][3][2], face tags [E][3] and refinement levels [E];
 * any pointer may be NULL */
dg_status dg_get_mesh(const dg_ctx* ctx, int32_t* n_elems, int32_t* vids, double* xyz, int32_t* bc, int32_t* levels);

/* Device-resident variants for callers that keep state on the GPU (torch tensors):
 * pointers are device pointers holding dg_ndof() doubles in the boundary layout. */
dg_status dg_set_state_device(dg_ctx* ctx, const double* dU);
dg_status dg_get_state_device(dg_ctx* ctx, double* dU);

/* Timing helper for benchmarks: replays nsteps of dg_advance on the context's
 * stream between two HIP events and returns the elapsed milliseconds. */
dg_status dg_time_advance(dg_ctx* ctx, int rk_scheme, double dt, int nsteps, float* ms_out);

/* ---- limiter and order selection (triangles; FP64 form) -------------------
 * The reference's limiter interface is DGLimiter::limit(DGCell*, time)
 * (DGLimiter.h:12), applied to every cell after each RK stage
 * (TimeIntegrator::limit, TimeIntegrator.cc:1414-1434; calls at :139-272,
 * :359-429, :850-987).  Its moment limiter (DGLimiter.cc:963-1133) indexes
 * quadrilateral tensor modes; this build limits triangles with a TVB minmod of
 * the linear part at the edge midpoints against the neighbour means (M: TVB
 * constant, 0 = TVD; nu: slope factor), dropping the modes of degree >= 2 of a
 * limited field.  Flags: one byte per element (caller order), bit f set when
 * field f was limited.  When enabled, dg_advance limits after every stage. */
dg_status dg_set_limiter(dg_ctx* ctx, int enabled, double tvb_m, double nu);
/* limit the current state once (TimeIntegrator::limit); flags_out may be NULL */
dg_status dg_limit(dg_ctx* ctx, uint8_t* flags_out);
/* flags of the last limiter pass (dg_limit or the last stage of dg_advance) */
dg_status dg_limiter_flags(dg_ctx* ctx, uint8_t* flags_out);

/* ---- non-physical states (DGCell::setToZero, DGCell.cc:251-279, which the
 * reference's assembly calls for a cell whose Euler state has rho < 0 or p < 0,
 * Euler2d::isPhysical, EulerLaw.cc:1398-1404, at a volume point,
 * DGCell.cc:198-223, or on one of its faces, DGCell.cc:639-732 with
 * parallel.cc:57-72 and TimeIntegrator.cc:1390-1405) --------------------------
 * dg_advance checks the state after its steps: DG_ERR_NONFINITE when a
 * coefficient is NaN or Inf; DG_ERR_NONPHYSICAL when (Euler, triangles) an
 * element is non-physical at a point its next residual evaluates (its volume
 * points and its traces at the points of each face rule), or when a recovery
 * pass met a non-physical mean.  With recovery enabled (Euler, triangles, FP64
 * form), every stage's output and the first stage's input pass through the
 * recovery: a non-physical element keeps its means and drops every mode >= 1,
 * as setToZero does.  All elements are checked on the same input state and each
 * rewrites only itself, so the result does not depend on any order (the
 * reference reduces cells one at a time inside its assembly). */
dg_status dg_set_recovery(dg_ctx* ctx, int enabled);
/* One pass over the current state: apply = 1 reduces the non-physical elements,
 * 0 only reports them.  flags_out (caller order, may be NULL): bit 0 a
 * non-physical point, bit 1 a non-physical mean, bit 2 a NaN or Inf coefficient. */
dg_status dg_recover(dg_ctx* ctx, int apply, uint8_t* flags_out);
/* the flags (as dg_recover's) OR-ed over the recovery passes of the last
 * dg_advance, or of the last dg_recover */
dg_status dg_recovery_flags(dg_ctx* ctx, uint8_t* flags_out);

/* Order selection (replaces DGAnalysis::adaptP / modifyP, DGPref.cc:24-60,
 * 602-630, whose face-jump bins choose orders 1..3): a modal-decay sensor on
 * field 0.  With top = energy of the degree-p modes and all = energy of every
 * mode of an element of order p, the next order is max(pmin, p-1) when
 * top > s_hi[p-1] all, min(pmax, p+1) when top < s_lo[p-1] all, else p
 * (elements with no energy above the mean keep p).  orders_out (caller order)
 * may be NULL; with apply the context re-buckets its elements by the new orders
 * and carries the state over (raised orders get zero modes, lowered orders drop
 * the top modes: DGCell::adaptOrder). */
typedef struct {
  double s_hi[4];         /* per current order 1..4 */
  double s_lo[4];
  int pmin, pmax;
} dg_sensor_opts;
dg_status dg_select_orders(dg_ctx* ctx, const dg_sensor_opts* opts, int apply, int32_t* orders_out);
/* current per-element orders, caller order */
dg_status dg_get_orders(dg_ctx* ctx, int32_t* orders_out);

/* ---- derivatives (no counterpart in the reference) ------------------------
 * Tangent-linear and adjoint residual products at a state U (boundary layout,
 * host buffers, FP64): dR = (dR/dU) dU and w = (dR/dU)^T z.  For callers that
 * differentiate the residual (tests, optimisation drivers). */
dg_status dg_residual_jvp(dg_ctx* ctx, const double* U, const double* dU, double* dR);
dg_status dg_residual_vjp(dg_ctx* ctx, const double* U, const double* z, double* w);

/* Discrete adjoint of nsteps RK4 steps (TimeIntegrator.cc:81-275) of the
 * unlimited residual, FP64 form.  dg_adjoint_forward advances the current state
 * nsteps steps of dt, keeping a checkpoint every ckpt_every steps on the device;
 * with a target state (boundary layout) it returns the objective
 * J = 1/2 sum_e |J_e| sum_k (rho_ek - rho*_ek)^2.  dg_adjoint_backward returns
 * dJ/dU(0) for the seed dJ/dU(T) (NULL: the gradient of that objective),
 * recomputing each segment's stage states from its checkpoint. */
dg_status dg_adjoint_forward(dg_ctx* ctx, double dt, int nsteps, int ckpt_every, const double* target, double* J_out);
dg_status dg_adjoint_backward(dg_ctx* ctx, const double* dJdUT, double* dJdU0);
/* device times of the last forward and backward calls (HIP events) */
dg_status dg_adjoint_timing(const dg_ctx* ctx, float* fwd_ms, float* bwd_ms);

/* ---- launch schedules and the tuner (no counterpart in the reference, whose
 * pthreads split is fixed by parallel.cc) ----------------------------------
 * A schedule is how the context's stage kernel is launched; except for the
 * matrix-core shape below it changes no state (every candidate computes each
 * coefficient with the same operations in the same order); the step residual
 * norm of dg_advance, a sum over workgroups and threads, follows the launch to
 * rounding.
 *   shape: the tiled triangle and tetrahedron kernels, 0 = 4 waves per tile of
 *          64 elements (256 threads), 1 = 8 waves (512 threads, same LDS);
 *          2 (the FP32 form of the Euler law) = 4 waves with the test
 *          contraction and the face lift on the matrix cores
 *          (v_mfma_f32_16x16x4_f32): its sums run in the matrix cores' order,
 *          so the state follows shape 0 to FP32 rounding, not bit for bit;
 *          Tetrahedra: 0 / 1 = 4 / 8 waves with the face traces over the
 *          coefficient rows in LDS, 2 / 3 / 4 = 8 / 16 / 4 waves with the rows
 *          kept (own traces from LDS, more LDS per tile);
 *          Mixed orders: 0 = one persistent launch over every order bucket,
 *          1 = one launch per order bucket (each with its order's registers
 *          and LDS), the buckets side by side on the context's side streams;
 *          compiled apart from shape 0, its states follow shape 0 to FP64
 *          rounding, not bit for bit;
 *   grid:  the persistent mixed-order launch (shape 0), workgroups walking the
 *          chunk list (0: one per chunk, at most 2048). */
typedef struct {
  int shape;
  int grid;
  int reserved[6];
} dg_schedule;
dg_status dg_get_schedule(const dg_ctx* ctx, dg_schedule* out);
/* DG_ERR_ARG for a schedule this context cannot run */
dg_status dg_set_schedule(dg_ctx* ctx, const dg_schedule* sched);

enum { DG_TUNE_FORCE = 1 };  /* time the candidates even when the cache has an entry */
#define DG_TUNE_MAX_CAND 16
typedef struct {
  dg_schedule best;
  int ncand;
  dg_schedule cand[DG_TUNE_MAX_CAND];
  float ms_per_step[DG_TUNE_MAX_CAND];  /* best of the timed repetitions, device time */
  int from_cache;                       /* 1: best read from the cache, nothing timed */
  char key[256];                        /* the configuration key of the cache */
} dg_tune_report;
/* Times each candidate schedule over `steps` steps of `scheme` with step dt on
 * this context's mesh and current state (saved before and restored after; the
 * limiter, when on, is part of every step), keeps the fastest and reports every
 * candidate.  cache (may be NULL): a text file of "key<TAB>shape grid" lines,
 * the key naming the device architecture and the configuration (dimension, law,
 * flux, orders, precision, elements, limiter, kernel); an entry for this key is
 * applied without timing unless flags has DG_TUNE_FORCE, and a timed result is
 * appended. */
dg_status dg_autotune(dg_ctx* ctx, int scheme, double dt, int steps, const char* cache, int flags,
                      dg_tune_report* report);

/* Measurement helper (no counterpart in the reference): one streaming copy of
 * n_doubles between two fresh device buffers; the reference point that scales
 * the HBM byte counters of the profiler.  Returns the kernel time. */
dg_status dg_copy_probe(int device, size_t n_doubles, float* ms_out);

/* Built-in exact fields: 0 Gaussian pulse carried by the scalar velocity on the
 * periodic unit square; 1 isentropic vortex, par = {beta, xc, yc, u0, v0, box
 * length}; 2 Sod tube (Examples.cc:62-83), par = {interface x}: the initial step
 * at t = 0 and the exact Riemann solution (Toro ch. 4) at t > 0; 3 PEC cavity mode. */
enum { DG_FIELD_GAUSSIAN = 0, DG_FIELD_VORTEX = 1, DG_FIELD_SOD = 2, DG_FIELD_CAVITY = 3 };

/* DGCell::L2ProjInitial of an exact field at time t over all cells
 * (DGCell.cc:1458-1483; volume rule of degree 2p); becomes the current state. */
dg_status dg_project(dg_ctx* ctx, int field, const double* par, double t);

/* The same projection of a caller's field: DGAnalysis's constructor projecting
 * the law's far-field FieldEvaluator into every cell (DGAnalysis.cc:200,
 * DGCell::L2ProjInitial).  fn(x, t, val, user) fills the Nf values at physical
 * point x on the host; a non-finite value is DG_ERR_ARG. */
dg_status dg_project_fn(dg_ctx* ctx, dg_field_fn fn, void* user, double t);

/* Per-field L2 norm of (state - exact field at t), rule of degree 2(p+1)
 * (DGCell::densityErrorL2, DGError.cc:805-831); err holds Nf doubles. */
dg_status dg_l2_error(dg_ctx* ctx, int field, const double* par, double t, double* err);

/* Structured triangle mesh (no counterpart in the reference, which reads Gmsh
 * files, mImportExport.cc:77-170): nx*ny squares on [x0,x0+lx]x[y0,y0+ly], each
 * split by the diagonal (i,j)-(i+1,j+1) into two counter-clockwise triangles,
 * elements row-major.  Interior lattice vertices move by a uniform +-jitter*h in
 * x and y drawn from splitmix64(seed ^ (2*vid)) and splitmix64(seed ^ (2*vid+1)).
 * side_tags: boundary tags of the bottom, right, top, left sides when not periodic.
 * Output buffers: vids [2*nx*ny*3], xyz [2*nx*ny*3*2], bc [2*nx*ny*3] (may be NULL). */
dg_status dg_mesh_tri_structured(int nx, int ny, double x0, double y0, double lx, double ly,
                                 int periodic_x, int periodic_y, double jitter, uint64_t seed,
                                 const int32_t side_tags[4], int32_t* vids, double* xyz, int32_t* bc);

/* Gmsh version-1 files (mImportExport::importGmshFile, mImportExport.cc:77-170;
 * periodic pairing as mMesh::setPeriodicBC, mMesh.cc:457-523): straight-sided
 * triangles (2D) or tetrahedra (3D, when the file holds any).  Boundary lines
 * (2D) or triangles (3D) carry physical tags onto the element faces they cover
 * (bc; untagged faces without neighbour: DG_TAG_FARFIELD; tag 100 interior).
 * Edges tagged 510 pair by equal y, 610 by equal x: the paired vertices get one
 * topological id, so the sides become face neighbours.  The vertex id of a
 * node (or of a class of paired nodes) is its smallest node number minus one.  Quadrilaterals are
 * DG_ERR_UNSUPPORTED unless split_quads (two triangles along the v0-v2 diagonal).
 * Call with vids = xyz = bc = NULL for the sizes in info, then with buffers of
 * [E][dim+1] ids, [E][dim+1][dim] coordinates, [E][dim+1] tags.  err_out (may be
 * NULL) receives a message on failure. */
typedef struct {
  int dim;
  int n_elems;
  int n_nodes;
  int n_quads_split;
  int n_periodic_pairs;
  int n_tagged;          /* boundary entities carrying a tag other than 100 / 510 / 610 */
} dg_gmsh_info;
dg_status dg_mesh_read_gmsh(const char* path, int split_quads, dg_gmsh_info* info, int32_t* vids, double* xyz,
                            int32_t* bc, char* err_out, size_t err_len);

/* The mesh half of DGAnalysis::write (mImportExport::exportDGFile,
 * mImportExport.cc:611-642): a Gmsh version-1 file of the descriptor that
 * dg_mesh_read_gmsh reads back to the same elements, faces and tags.  One node
 * per (vertex id, coordinates); faces without neighbour become lines (2D) or
 * triangles (3D) carrying their bc tag; the two sides of a periodic face pair
 * (one vertex id at two positions, a translation along x or y) become lines
 * tagged 510 (translated along x: paired by equal y) or 610 (along y).
 * Periodic tetrahedral meshes are DG_ERR_UNSUPPORTED. */
dg_status dg_mesh_write_gmsh(const dg_mesh_desc* mesh, const char* path, char* err_out, size_t err_len);

/* ---- checkpoint / restart (DGAnalysis::write / read, DGPref.cc:712-768) ----
 * The data stream of DGAnalysis::write: t, the iteration, "dim E F", then per
 * element in caller order its mMesh::writeMEntity record (entity type 2 or 4
 * and the vertex ids, mMesh.cc:401-412, written as node numbers: vertex id + 1,
 * the reference's Vertex::getId() of a Gmsh node) followed by DGCell::write (the order and
 * the Nf*Np coefficients field by field, DGCell.cc:1523-1530), then one record
 * per face (type 1 or 2 and its vertex ids).  Coefficients are written with 17
 * significant digits (the reference: 16), so a restart holds the same doubles. */
dg_status dg_write_restart(dg_ctx* ctx, const char* path, double t, int iter);
/* Reads such a stream (DGAnalysis::read, DGPref.cc:742-768, which expects a
 * geometric order after each entity type that write never writes; this reader
 * takes the records as write produces them).  Cell records are matched to the
 * context's elements by their vertex ids (node number - 1; any record order;
 * each element's vertices must come in the context's order), so a file the
 * reference wrote over a Gmsh mesh without periodic sides reads into a context
 * built from that mesh file.  On periodic sides the reference records the node
 * number of each image while dg_mesh_read_gmsh gives a pair one id: those
 * records find no element (DG_ERR_ARG names the record).  A triangle context in the FP64
 * form takes the orders of the file (re-bucketing as dg_select_orders does);
 * otherwise they must equal the context's.  Face records are not needed. */
dg_status dg_read_restart(dg_ctx* ctx, const char* path, double* t_out, int* iter_out);

/* Norms and conserved totals (the sums of DGPref.cc:249-394 and the total mass
 * of DGAnalysis.cc:554, DGCell::computeMass, DGError.cc:833-845), per field:
 *   totals[f] = sum_e int_K u_f and energy[f] = sum_e int_K u_f^2 (device
 *   reduction over the state, summed in a fixed order);
 *   l1, l2, linf of (u_f - exact_f) at t against a built-in exact field
 *   (DG_FIELD_*, parameters exact_par; DGCell::L1exact, densityErrorL2,
 *   LinfError, DGError.cc:549-580, 805-831, which take field 0), at the points
 *   of the rule of degree 2(p+1).
 * Every output (Nf doubles) may be NULL; error norms need exact_field >= 0. */
dg_status dg_norms(dg_ctx* ctx, double t, int exact_field, const double* exact_par, double* l1, double* l2,
                   double* linf, double* totals, double* energy);

/* ---- state export (DGAnalysis::exportGmsh, DGSensor.cc:337-525, and
 * exportGmsh_Volume, :800-) ----------------------------------------------------
 * Lattice of a reference simplex split nsub times: points (i, j[, k]) / nsub
 * with i + j [+ k] <= nsub, k outer, i inner.  dg_lattice_points gives their
 * number (0 for nsub outside 1..8).  dg_sample_lattice evaluates every field
 * at every element's lattice points on the device: out[E][Nf][points], caller
 * element order, n = E * Nf * points. */
size_t dg_lattice_points(int dim, int nsub);
dg_status dg_sample_lattice(dg_ctx* ctx, int nsub, double* out, size_t n);
/* Gmsh post-processing views, one file "<prefix>-sample<iteration><name>.pos"
 * per field of interest (ST/VT triangles, SS/VS tetrahedra):
 *   Euler:   velocity, pressure, mach, density, error (rho u - far field rho u),
 *            entropy (against the far-field state), limit (Euler::getNameAndSize,
 *            Euler2d::getIthFieldOfInterest, EulerLaw.cc:1240-1340);
 *   Scalar:  error and exact (with an exact field), numEr, function, limit
 *            (ScalarLaw.cc:111-165);
 *   Maxwell2d: Intensity_RefCond (Ez^2), HxField_RefCond, HyField_RefCond,
 *            EField_RefCond (MaxwellLaw.cc:256-330);
 *   Burgers2D: function, error and exact (with an exact field), numEr
 *            (Burgers2D.cc:121-200);
 *   Maxwell (3D): E and H vectors, energy density, error |E - E_exact| (with an
 *            exact field);
 *   all:     order.
 * "limit" is the element's flag byte of the last limiter pass.  nsub <= 0: each
 * triangle split by the highest element order (the reference: its own order),
 * each tetrahedron into 1 (order 1) or 8 (red refinement).  exact_field:
 * DG_FIELD_* at time t with exact_par, or -1. */
dg_status dg_export_pos(dg_ctx* ctx, const char* prefix, int iteration, int nsub, double t, int exact_field,
                        const double* exact_par);

/* Kuhn subdivision of an nx x ny x nz lattice of boxes on [0,lx]x[0,ly]x[0,lz]
 * into 6 tetrahedra per box sharing its main diagonal (faces match across boxes).
 * vids[E][4], xyz[E][4][3], E = 6 nx ny nz.  With the Maxwell law every face
 * without a neighbour is a perfect electric conductor. */
dg_status dg_mesh_tet_structured(int nx, int ny, int nz, double lx, double ly, double lz, int32_t* vids,
                                 double* xyz);

#ifdef __cplusplus
}
#endif
#endif
