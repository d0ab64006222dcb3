// Host C++ mirror of the reference's call surface for the DG hot path:
//   ConservationLaw and its Euler / Scalar / Maxwell / Burgers subclasses
//     (ConservationLaw.h:12-60, EulerLaw.h, ScalarLaw.h, MaxwellLaw.h, Burgers.h),
//   DGAnalysis (DGAnalysis.h:110-135),
//   TimeIntegrator::advanceInTime and its RK subclasses (TimeIntegrator.h:8-91).
// The same method names and argument meanings; the work is done on the GPU
// through the C ABI (include/dg_capi.h).  The law classes keep their pointwise
// methods (Fi, riemannSolver, boundaryFlux, maximumEigenValue, isPhysical)
// evaluated on the host with the same arithmetic the kernels use, and describe
// themselves to the device as a dg_law_desc -- there is no virtual dispatch on
// the device.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dg_capi.h"

namespace dgpar {

// point and vector in the reference's sense (mPoint.h, mVector.h)
struct mPoint {
  double p[3] = {0, 0, 0};
  mPoint() = default;
  mPoint(double x, double y, double z = 0) : p{x, y, z} {}
  double operator()(int i) const { return p[i]; }
  double& operator()(int i) { return p[i]; }
};

struct mVector {
  double v[3] = {0, 0, 0};
  mVector() = default;
  mVector(double x, double y, double z = 0) : v{x, y, z} {}
  double operator()(int i) const { return v[i]; }
  double& operator()(int i) { return v[i]; }
  double operator*(const mVector& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
};

// FieldEvaluator.h:9-14
class FieldEvaluator {
 public:
  virtual ~FieldEvaluator() = default;
  virtual void eval(const mPoint& space, double time, double* val) const = 0;
  virtual int order() const { return 0; }
};

class ConstantField : public FieldEvaluator {
  std::vector<double> v_;

 public:
  explicit ConstantField(std::vector<double> v) : v_(std::move(v)) {}
  void eval(const mPoint&, double, double* val) const override {
    for (size_t i = 0; i < v_.size(); i++) val[i] = v_[i];
  }
};

class DGError : public std::runtime_error {
 public:
  dg_status status;
  DGError(dg_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// ---------------------------------------------------------------------------
class ConservationLaw {
 protected:
  FieldEvaluator* farField = nullptr;
  FieldEvaluator* exactField = nullptr;

 public:
  virtual ~ConservationLaw() = default;
  FieldEvaluator* getFarField() { return farField; }
  FieldEvaluator* getExactSolution() { return exactField; }
  virtual int getNbFields() const = 0;
  virtual void Fi(const mPoint& position, double* Q, mVector* flux) const = 0;
  virtual void RHS(double*, const mPoint&, double, double* rhs) const {
    for (int i = 0; i < getNbFields(); i++) rhs[i] = 0.0;
  }
  virtual void riemannSolver(mVector& n, const mPoint& position, double* left, double* right, double* flux) const = 0;
  virtual void boundaryFlux(mVector& n, int BoundaryId, const mPoint& position, double* field, double* flux,
                            double T) const = 0;
  virtual double maximumEigenValue(const double* field, const mPoint& position) const = 0;
  virtual bool isPhysical(double* val) const { return true; }
  // the device-side description of this law (no counterpart in the reference)
  virtual dg_law_desc describe() const = 0;
  // a velocity FieldEvaluator the device evaluates at its quadrature points
  // (ScalarLaw's velocityField), or null
  virtual FieldEvaluator* velocityField() const { return nullptr; }
};

// ScalarLaw (ScalarLaw.cc:19-97): the reference's constructor takes the velocity
// as a FieldEvaluator (ScalarLaw(farField, velocityField, exactField)); the
// second constructor takes a constant velocity
class ScalarLaw : public ConservationLaw {
  double vel_[3];
  FieldEvaluator* velField_ = nullptr;
  void velocity(const mPoint& p, double* v) const;

 public:
  ScalarLaw(FieldEvaluator* far, FieldEvaluator* velocity, FieldEvaluator* exact);
  ScalarLaw(FieldEvaluator* far, const double vel[3], FieldEvaluator* exact = nullptr);
  FieldEvaluator* velocityField() const override { return velField_; }
  int getNbFields() const override { return 1; }
  void Fi(const mPoint&, double* Q, mVector* flux) const override;
  void riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const override;
  void boundaryFlux(mVector& n, int, const mPoint&, double* Q, double* flux, double) const override;
  double maximumEigenValue(const double*, const mPoint&) const override;
  dg_law_desc describe() const override;
};

// Maxwell2d (MaxwellLaw.cc:21-200): (Hx, Hy, Ez), its own upwind interface
// state, walls (40000, 20000) mirror H, other boundaries take the far field
class Maxwell2d : public ConservationLaw {
 public:
  explicit Maxwell2d(FieldEvaluator* far) { farField = far; exactField = far; }
  int getNbFields() const override { return 3; }
  void Fi(const mPoint&, double* Q, mVector* flux) const override;
  void riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const override;
  void boundaryFlux(mVector& n, int BoundaryId, const mPoint&, double* Q, double* flux, double T) const override;
  double maximumEigenValue(const double*, const mPoint&) const override { return 1.0; }
  dg_law_desc describe() const override;
};

// Burgers2D (Burgers2D.cc:22-92): F = (u^2/2, u^2/2), central flux, far field
// upwind on boundaries
class Burgers2D : public ConservationLaw {
 public:
  explicit Burgers2D(FieldEvaluator* far) { farField = far; exactField = far; }
  int getNbFields() const override { return 1; }
  void Fi(const mPoint&, double* Q, mVector* flux) const override;
  void riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const override;
  void boundaryFlux(mVector& n, int BoundaryId, const mPoint&, double* Q, double* flux, double T) const override;
  double maximumEigenValue(const double*, const mPoint&) const override { return 1.0; }
  dg_law_desc describe() const override;
};

// Euler2d (EulerLaw.cc:47-69, 182-296, 1012-1034, 1398-1404) with its own
// numerical flux, the exact Godunov flux (EulerLaw.cc:602-887); Roe in
// Euler2dRoe, local Lax-Friedrichs (new) in Euler2dLF
class Euler2d : public ConservationLaw {
 protected:
  double gamma;
  int flux_kind = DG_FLUX_EXACT;

 public:
  Euler2d(FieldEvaluator* far, double Gamma = 1.4);
  int getNbFields() const override { return 4; }
  void Fi(const mPoint&, double* Q, mVector* flux) const override;
  void riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const override;
  void boundaryFlux(mVector& n, int BoundaryId, const mPoint&, double* Q, double* flux, double T) const override;
  double maximumEigenValue(const double* Q, const mPoint&) const override;
  bool isPhysical(double* Q) const override;
  dg_law_desc describe() const override;
};

class Euler2dLF : public Euler2d {
 public:
  Euler2dLF(FieldEvaluator* far, double Gamma = 1.4) : Euler2d(far, Gamma) { flux_kind = DG_FLUX_LF; }
};

class Euler2dRoe : public Euler2d {
 public:
  // entropy_fix: Harten's fix (delta = 0.2 c), which the reference's Roe flux lacks
  Euler2dRoe(FieldEvaluator* far, double Gamma = 1.4, bool entropy_fix = false) : Euler2d(far, Gamma) {
    flux_kind = entropy_fix ? DG_FLUX_ROE_EF : DG_FLUX_ROE;
  }
};

// Euler2d with a second far-field state on faces tagged DG_TAG_FARFIELD2 (the
// reference evaluates its far-field FieldEvaluator at the boundary point; a shock
// tube's two ends hold the two initial states)
class Euler2dTwoStates : public Euler2dRoe {
  double second_[4];

 public:
  Euler2dTwoStates(FieldEvaluator* far, const double second[4], double Gamma = 1.4, bool entropy_fix = true)
      : Euler2dRoe(far, Gamma, entropy_fix) {
    for (int i = 0; i < 4; i++) second_[i] = second[i];
  }
  dg_law_desc describe() const override;
};

// Maxwell's equations in 3D with eps = mu = 1 (the reference's MaxwellLaw.cc:21-200
// is the 2D TE form): Q = (E, H), F.n = (-n x H, n x E), upwind or central flux,
// perfect-conductor walls (E -> -E, H -> H in the ghost)
class Maxwell3d : public ConservationLaw {
  int flux_kind;

 public:
  explicit Maxwell3d(bool upwind = true) : flux_kind(upwind ? DG_FLUX_UPWIND : DG_FLUX_CENTRAL) {}
  int getNbFields() const override { return 6; }
  void Fi(const mPoint&, double* Q, mVector* flux) const override;
  void riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const override;
  void boundaryFlux(mVector& n, int BoundaryId, const mPoint&, double* Q, double* flux, double T) const override;
  double maximumEigenValue(const double*, const mPoint&) const override { return 1.0; }
  dg_law_desc describe() const override;
};

// ---------------------------------------------------------------------------
// DGLimiter (DGLimiter.h:9-14): the reference limits one cell at a time from the
// integrator; here the GPU limits every cell after every stage, so the plugin
// carries its parameters to the context.
class DGLimiter {
 public:
  virtual ~DGLimiter() = default;
  virtual double tvbM() const = 0;
  virtual double nu() const = 0;
};
// triangle TVB limiter with positivity scaling (new design: DGMomentLimiter,
// DGLimiter.cc:963-1133, indexes quadrilateral modes only)
class DGTriangleLimiter : public DGLimiter {
  double m_, nu_;

 public:
  explicit DGTriangleLimiter(double tvb_m = 0.0, double nu = 1.0) : m_(tvb_m), nu_(nu) {}
  double tvbM() const override { return m_; }
  double nu() const override { return nu_; }
};

// ---------------------------------------------------------------------------
// A mesh handed to DGAnalysis (the path's view of mDGMesh)
struct Mesh {
  int dim = 2;
  std::vector<int32_t> vids, bc, order;
  std::vector<double> xyz;
  int n_elems() const { return (int)(vids.size() / (dim + 1)); }
  static Mesh structuredTri(int nx, int ny, double x0, double y0, double lx, double ly, bool periodic_x,
                            bool periodic_y, double jitter, uint64_t seed, const int32_t side_tags[4]);
  // Kuhn subdivision of an n^3 lattice of boxes (6 tetrahedra per box)
  static Mesh structuredTet(int nx, int ny, int nz, double lx, double ly, double lz);
  // a Gmsh version-1 file (mImportExport::importGmshFile): straight triangles or
  // tetrahedra, boundary tags, 510/610 periodic sides joined
  static Mesh readGmsh(const char* path, bool split_quads = false);
  // the mesh as a Gmsh version-1 file that readGmsh reads back
  // (mImportExport::exportDGFile, mImportExport.cc:611-642)
  void writeGmsh(const char* path) const;
  dg_mesh_desc describe() const;
};

class DGAnalysis;

// TimeIntegrator (TimeIntegrator.h:8-29): advanceInTime returns the residual norm
class TimeIntegrator {
 protected:
  DGAnalysis* an_;
  int scheme_;

 public:
  TimeIntegrator(DGAnalysis* a, int scheme) : an_(a), scheme_(scheme) {}
  virtual ~TimeIntegrator() = default;
  virtual double advanceInTime(double t, double dt);
  int scheme() const { return scheme_; }
};
class ForwardEuler : public TimeIntegrator {
 public:
  explicit ForwardEuler(DGAnalysis* a) : TimeIntegrator(a, DG_RK_FE) {}
};
class RungeKuttaTVD2 : public TimeIntegrator {
 public:
  explicit RungeKuttaTVD2(DGAnalysis* a) : TimeIntegrator(a, DG_RK_TVD2) {}
};
class RungeKuttaTVD3 : public TimeIntegrator {
 public:
  explicit RungeKuttaTVD3(DGAnalysis* a) : TimeIntegrator(a, DG_RK_TVD3) {}
};
class RungeKutta4 : public TimeIntegrator {
 public:
  explicit RungeKutta4(DGAnalysis* a) : TimeIntegrator(a, DG_RK4) {}
};
// SSP-RK2 with local time steps per size level (TimeIntegrator.h:45-70,
// TimeIntegrator.cc:442-833): dt is the coarsest level's step; advanceInTime
// returns 0 as the reference's does
class RungeKuttaTVD2adaptive : public TimeIntegrator {
 public:
  explicit RungeKuttaTVD2adaptive(DGAnalysis* a) : TimeIntegrator(a, DG_RK_TVD2_LTS) {}
};

// DGAnalysis (DGAnalysis.h:110-135): owns the GPU context and the integrator,
// chosen from the order as the reference does (DGAnalysis.cc:239-242).
class DGAnalysis {
  dg_ctx* ctx_ = nullptr;
  Mesh mesh_;
  ConservationLaw* law_;
  TimeIntegrator* integrator_ = nullptr;
  int order_;
  int n_elems_ = 0;

 public:
  double TACT = 0.0, TEND = 0.2, CFLMAX = 0.5;
  int ITER = 0;

  // limiter: NULL (no limiting) or a DGLimiter applied after every RK stage
  DGAnalysis(const Mesh& mesh, ConservationLaw* law, DGLimiter* limiter, int order, int precision = DG_PREC_FP64);
  ~DGAnalysis();
  DGAnalysis(const DGAnalysis&) = delete;
  DGAnalysis& operator=(const DGAnalysis&) = delete;

  dg_ctx* context() { return ctx_; }
  TimeIntegrator* integrator() { return integrator_; }
  size_t ndof() const { return dg_ndof(ctx_); }
  void check(dg_status s) const;

  void setState(const std::vector<double>& U);
  std::vector<double> getState() const;
  // L2 projection of a built-in exact field (DGCell::L2ProjInitial)
  void project(int field, const double* par, double t = 0.0);
  std::vector<double> residual(double t = 0.0);
  double computeDt(double cfl);
  // limit every cell once (TimeIntegrator::limit); per-element flags, caller order
  std::vector<uint8_t> limit();
  // p-adaptation (DGAnalysis::adaptP, DGPref.cc:602-630) with the modal-decay
  // sensor; re-buckets the elements by their new orders; returns the orders
  std::vector<int32_t> adaptP(const dg_sensor_opts& opts);
  std::vector<int32_t> orders() const;
  // local time stepping (DGAnalysis::sortCellsBySize, DGAnalysis.cc:749-890, and
  // the integrator the reference selects with it, DGAnalysis.cc:232-234): cells
  // binned by size into at most max_levels levels, the integrator replaced by
  // RungeKuttaTVD2adaptive; computeDt then returns the coarsest level's step.
  // Returns each cell's level (caller order).
  std::vector<int32_t> sortCellsBySize(int max_levels = 6);
  // h-adaptivity (DGAnalysis::adaptH's mesh operations, DGPref.cc:78-192): cells
  // marked > 0 are split into four (at most max_level times, LEVEL_OF_REFINEMENT),
  // families of four children marked < 0 merged, the state L2-projected; the
  // cells then follow the new mesh (refinedMesh()).  Returns the new cell count.
  int adaptH(const std::vector<int8_t>& marks, int max_level = 1);
  // the current mesh after adaptH (the reference exports it as refined-N.msh)
  Mesh refinedMesh() const;
  // launch schedule of the stage kernel (no counterpart in the reference): the
  // tuner times every candidate over `steps` steps of the integrator's scheme
  // with step dt on the current state (restored after) and keeps the fastest;
  // cache: optional text file of tuned entries per device and configuration
  dg_tune_report autotune(double dt, int steps = 5, const char* cache = nullptr, bool force = false);
  dg_schedule schedule() const;
  // checkpoint / restart (DGAnalysis::write / read, DGPref.cc:712-768): the mesh
  // as a Gmsh file and the data stream (TACT, ITER, orders, coefficients); read
  // takes the data stream of a context over the same mesh and sets TACT, ITER
  void write(const char* mesh_path, const char* data_path) const;
  void read(const char* data_path);
  // Gmsh post-processing views (DGAnalysis::exportGmsh, DGSensor.cc:337-525):
  // one <fileName>-sample<iteration><field>.pos file per field of interest
  void exportGmsh(const char* fileName, int iteration) const;
  // the reference's time loop (commented out in DGAnalysis.cc:480-731): dt from the
  // CFL rule, N = ceil((TEND-TACT)/dt) steps of dt = (TEND-TACT)/N; returns N
  int run();
};

}  // namespace dgpar
