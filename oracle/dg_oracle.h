/*
 * dg_oracle -- CPU FP64 restatement of DG-Parallel's explicit DG residual and
 * Runge-Kutta step.  TEST INFRASTRUCTURE ONLY: it is the parity checker and the
 * reported CPU baseline.  Nothing in the product links, loads or calls it; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every routine names the reference file:line it restates.  Parity status of
 * each function (what pins it) is listed in the header of dg_oracle.c.
 *
 * State layout (both input and output): element-major, then field, then mode,
 * i.e. U[e][f][k] with Np(e) modes -- the reference's per-cell coefficient
 * block `i*FS+j` (DGCell.h:31), concatenated over elements in input order.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* OR_LAW_MAXWELL2D: Maxwell2d (MaxwellLaw.cc:21-200); OR_LAW_BURGERS2D: Burgers2D (Burgers2D.cc:22-92) */
enum { OR_LAW_SCALAR = 0, OR_LAW_EULER2D = 1, OR_LAW_MAXWELL3D = 2, OR_LAW_MAXWELL2D = 3, OR_LAW_BURGERS2D = 4 };
/* OR_FLUX_EXACT: Euler2d's own riemannSolver, the exact Godunov flux (EulerLaw.cc:602-887) */
enum { OR_FLUX_UPWIND = 0, OR_FLUX_LF = 1, OR_FLUX_ROE = 2, OR_FLUX_CENTRAL = 3, OR_FLUX_ROE_EF = 4, OR_FLUX_EXACT = 5 };
enum { OR_RK_FE = 0, OR_RK_TVD2 = 1, OR_RK_TVD3 = 2, OR_RK4 = 3 };

typedef struct {
  int dim;                 /* 2: triangles, 3: tetrahedra */
  int n_elems;
  const int32_t* vids;     /* [E][dim+1] topological vertex ids (periodic images share an id) */
  const double* xyz;       /* [E][dim+1][dim] corner coordinates of each element */
  const int32_t* bc;       /* [E][dim+1] tag for faces without a neighbour (NULL: 0) */
  const int32_t* order;    /* [E] polynomial order of each element */
} or_mesh;

typedef struct {
  int law;                 /* OR_LAW_* */
  int flux;                /* OR_FLUX_* */
  double gamma;            /* Euler */
  double vel[3];           /* scalar advection velocity */
  double farfield[8];      /* ghost state for boundary faces with tag != 40000 */
} or_law;

typedef struct or_ctx or_ctx;

/* a FieldEvaluator (FieldEvaluator.h:12): val at the physical point x and time t */
typedef void (*or_field_fn)(const double* x, double t, double* val, void* user);

/* builds connectivity, geometry and quadrature tables; NULL on bad input */
or_ctx* or_create(const or_mesh* m, const or_law* law);
void or_destroy(or_ctx* c);
int or_nfields(const or_ctx* c);
long or_ndof(const or_ctx* c);                 /* sum_e Nf*Np(e) */
int or_nfaces(const or_ctx* c);
/* faces as (elemL, localL, elemR, localR); elemR = -1 on boundary faces;
 * elemL < elemR for interior faces */
void or_faces(const or_ctx* c, int32_t* out4);

/* R = int F.grad(phi) + S phi - oint Fhat phi  (DGCell.cc:184-249, 639-732) */
int or_residual(or_ctx* c, const double* U, double t, double* R, int nthreads);
/* ScalarLaw's velocity FieldEvaluator (ScalarLaw.cc:29-97): the flux, the
 * Riemann and boundary fluxes and the time step evaluate it at their point;
 * face fluxes at the face's point in the left cell's frame (0: scalar law only) */
int or_set_velocity_field(or_ctx* c, or_field_fn fn, void* user);
/* the same loops summing the magnitudes of the terms, |volume term| per point
 * and |face term| per face point: the scale of the residual's summation, which
 * bounds its rounding error (test hook) */
int or_residual_abs(or_ctx* c, const double* U, double t, double* Rabs, int nthreads);
/* one step of the chosen scheme in place; returns the residual norm the
 * reference's advanceInTime returns (TimeIntegrator.cc:81-275, 279-438, 838-990, 993-1048) */
double or_step(or_ctx* c, int scheme, double* U, double t, double dt, int nthreads);
/* global min of CFL*size/((2p+1) lambda) (DGCell.cc:1155-1262) */
double or_dt(or_ctx* c, const double* U, double cfl);
/* local time stepping (triangles; new design over the size bins of
 * DGAnalysis::sortCellsBySize, DGAnalysis.cc:749-890): level of each element
 * (bin b: size <= 1.0001 smax / 2^b, the largest b < max_levels; face
 * neighbours raised to within one level), returns the number of levels;
 * or_lts_dt: min over elements of their or_dt candidate x 2^level; or_lts_step:
 * one step of size h in place, level-b elements taking 2^b Heun substeps inside
 * one shared sequence of 2^levels stages (every stage evaluated everywhere) */
int or_lts_levels(or_ctx* c, int max_levels, int32_t* level);
double or_lts_dt(or_ctx* c, const double* U, double cfl, const int32_t* level);
int or_lts_step(or_ctx* c, const int32_t* level, double* U, double t, double h, int nthreads);
/* per-element detJ (reference-to-physical) and mean of field f */
void or_detj(const or_ctx* c, double* out);

/* basis values/gradients of order p on the reference simplex, from this
 * file's own long-double Gram-Schmidt (ShapeFunctionList.cc:120-175) */
int or_basis(int dim, int p, const double* uvw, double* phi, double* dphi);

/* L2 projection of a built-in exact field (DGCell.cc:1458-1483):
 * field 0: Gaussian pulse translated by vel*t (periodic unit square)
 * field 1: isentropic vortex (gamma, beta=par[0], centre par[1],par[2], free stream
 *          (1, par[3], par[4], 1), periodic box [0,par[5]]^2) at time t
 * field 2: Sod tube, interface x = par[0]: the initial step at t = 0, the exact
 *          Riemann solution (Toro ch. 4) at t > 0
 * field 3: PEC cavity mode (1,1,1) with E amplitude (1,-1,0) at time t */
int or_project(or_ctx* c, int field, const double* par, double t, double* U);
/* the same projection of a caller's field fn(x, t, val, user) (the reference's
 * DGAnalysis projecting its far-field FieldEvaluator, DGAnalysis.cc:200) */
int or_project_fn(or_ctx* c, or_field_fn fn, void* user, double t, double* U);
/* evaluates the same exact fields at a point */
int or_exact(int field, const double* par, const or_law* law, const double* x, double t, double* val);
/* test hooks: the law's numerical flux Fhat(uL, uR).n through one face point
 * (riemannSolver); the exact solution of a 1D Riemann problem with primitive
 * states wl, wr = (rho, u, p) sampled at s = x/t */
int or_riemann(const or_law* law, const double* n, const double* uL, const double* uR, double* out);
int or_riemann_exact(double gamma, const double* wl, const double* wr, double s, double* out);
/* per-field L2 error of U against an exact field, rule of degree 2(p+1)
 * (DGError.cc:805-831) */
int or_l2_error(or_ctx* c, const double* U, int field, const double* par, double t, double* err);

/* limiter (triangles, new design; the reference's DGMomentLimiter is
 * quad-only): limits U in place, flags[e] bit f set when field f was limited
 * (flags may be NULL).  or_set_limiter makes or_step limit after every stage. */
int or_limit(or_ctx* c, double* U, double tvb_m, double nu, uint8_t* flags);
void or_set_limiter(or_ctx* c, int on, double tvb_m, double nu);
/* non-physical recovery (DGCell::setToZero, DGCell.cc:251-279): elements whose
 * Euler state has rho < 0 or p < 0 at a volume point or at a point of one of
 * their faces keep their means and drop modes >= 1, in place; flags[e] (may be
 * NULL): bit 0 reduced, bit 1 with a non-physical mean, bit 2 a NaN or Inf
 * coefficient; returns the number reduced.
 * or_set_recovery makes or_step apply it to every stage's residual input. */
int or_recover(or_ctx* c, double* U, uint8_t* flags);
void or_set_recovery(or_ctx* c, int on);
/* modal-decay sensor: next order per element (s_hi, s_lo indexed by order-1) */
int or_sensor(or_ctx* c, const double* U, const double* s_hi, const double* s_lo, int pmin, int pmax,
              int32_t* orders);

/* test hooks: substitute the triangle basis coefficient matrix (n x n, row i =
 * coefficients of phi_i over the Pascal monomials) or the triangle rule of one
 * degree ({u, v, weight} with weights summing to 1); or_reset_tables() restores
 * the oracle's own.  Affects contexts created afterwards. */
int or_set_tri_basis(int n, const double* C);
int or_set_tri_rule(int degree, int n, const double* pts);
void or_reset_tables(void);

#ifdef __cplusplus
}
#endif
#endif
