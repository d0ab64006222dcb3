// Host C++ mirror of the reference's ConservationLaw / DGAnalysis /
// TimeIntegrator surface over the C ABI.
#include "dg_host.h"

#include <algorithm>

#include "../csrc/dg_laws.h"

namespace dgpar {

static dgp::LawParams params_of(const dg_law_desc& d) {
  dgp::LawParams L{};
  L.gamma = d.gamma;
  for (int i = 0; i < 3; i++) L.vel[i] = d.vel[i];
  for (int i = 0; i < 8; i++) L.farfield[i] = d.farfield[i];
  L.flux = d.flux;
  return L;
}

static void far_state(FieldEvaluator* f, int n, double* out) {
  if (f) {
    f->eval(mPoint(0, 0, 0), 0.0, out);
  } else {
    for (int i = 0; i < n; i++) out[i] = 0.0;
  }
}

// ---------------------------------------------------------------------------
ScalarLaw::ScalarLaw(FieldEvaluator* far, const double vel[3], FieldEvaluator* exact) {
  farField = far;
  exactField = exact;
  for (int i = 0; i < 3; i++) vel_[i] = vel[i];
}

void ScalarLaw::Fi(const mPoint&, double* Q, mVector* flux) const {
  for (int d = 0; d < 3; d++) flux[0](d) = vel_[d] * Q[0];
}

void ScalarLaw::riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const {
  dgp::LawParams L = params_of(describe());
  dgp::ScalarLawD::riemann<double>(L, n.v, uL, uR, flux);
}

void ScalarLaw::boundaryFlux(mVector& n, int, const mPoint&, double* Q, double* flux, double) const {
  dgp::LawParams L = params_of(describe());
  dgp::ScalarLawD::boundary<double>(L, dgp::BC_FARFIELD, n.v, Q, flux);
}

double ScalarLaw::maximumEigenValue(const double*, const mPoint&) const {
  return std::fabs(vel_[0]) + std::fabs(vel_[1]) + std::fabs(vel_[2]);
}

dg_law_desc ScalarLaw::describe() const {
  dg_law_desc d{};
  d.law = DG_LAW_SCALAR;
  d.flux = DG_FLUX_UPWIND;
  for (int i = 0; i < 3; i++) d.vel[i] = vel_[i];
  far_state(farField, 1, d.farfield);
  return d;
}

// ---------------------------------------------------------------------------
Euler2d::Euler2d(FieldEvaluator* far, double Gamma) : gamma(Gamma) {
  farField = far;
  exactField = far;
}

void Euler2d::Fi(const mPoint&, double* Q, mVector* flux) const {
  dgp::LawParams L = params_of(describe());
  double F[4][2];
  dgp::EulerLawD::flux<double>(L, Q, F);
  for (int f = 0; f < 4; f++) {
    flux[f](0) = F[f][0];
    flux[f](1) = F[f][1];
    flux[f](2) = 0.0;
  }
}

void Euler2d::riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const {
  dgp::LawParams L = params_of(describe());
  dgp::EulerLawD::riemann<double>(L, n.v, uL, uR, flux);
}

void Euler2d::boundaryFlux(mVector& n, int BoundaryId, const mPoint&, double* Q, double* flux, double) const {
  if (BoundaryId == DG_TAG_CHARACTERISTIC) throw DGError(DG_ERR_UNSUPPORTED, "characteristic far field not built");
  dgp::LawParams L = params_of(describe());
  dgp::EulerLawD::boundary<double>(L, BoundaryId == DG_TAG_WALL ? dgp::BC_WALL : dgp::BC_FARFIELD, n.v, Q, flux);
}

double Euler2d::maximumEigenValue(const double* Q, const mPoint&) const {
  dgp::LawParams L = params_of(describe());
  return dgp::EulerLawD::maxev<double>(L, Q);
}

bool Euler2d::isPhysical(double* Q) const {
  dgp::LawParams L = params_of(describe());
  return dgp::EulerLawD::physical<double>(L, Q);
}

dg_law_desc Euler2d::describe() const {
  dg_law_desc d{};
  d.law = DG_LAW_EULER2D;
  d.flux = flux_kind;
  d.gamma = gamma;
  far_state(farField, 4, d.farfield);
  return d;
}

// ---------------------------------------------------------------------------
Mesh Mesh::structuredTri(int nx, int ny, double x0, double y0, double lx, double ly, bool periodic_x,
                         bool periodic_y, double jitter, uint64_t seed, const int32_t side_tags[4]) {
  Mesh m;
  m.dim = 2;
  size_t E = (size_t)2 * nx * ny;
  m.vids.resize(E * 3);
  m.xyz.resize(E * 6);
  m.bc.resize(E * 3);
  dg_status s = dg_mesh_tri_structured(nx, ny, x0, y0, lx, ly, periodic_x, periodic_y, jitter, seed, side_tags,
                                       m.vids.data(), m.xyz.data(), m.bc.data());
  if (s != DG_OK) throw DGError(s, "dg_mesh_tri_structured rejected its arguments");
  return m;
}

// ---------------------------------------------------------------------------
double TimeIntegrator::advanceInTime(double t, double dt) {
  double r = 0.0;
  an_->check(dg_advance(an_->context(), scheme_, t, dt, 1, &r));
  return r;
}

DGAnalysis::DGAnalysis(const Mesh& mesh, ConservationLaw* law, void* limiter, int order, int precision)
    : law_(law), order_(order) {
  if (limiter) throw DGError(DG_ERR_UNSUPPORTED, "limiters are not built in this entry point");
  dg_mesh_desc md{};
  md.dim = mesh.dim;
  md.n_elems = mesh.n_elems();
  md.vids = mesh.vids.data();
  md.xyz = mesh.xyz.data();
  md.bc = mesh.bc.empty() ? nullptr : mesh.bc.data();
  md.order = mesh.order.empty() ? nullptr : mesh.order.data();
  dg_law_desc ld = law->describe();
  dg_options opt{};
  opt.order = order;
  opt.precision = precision;
  opt.use_graph = 1;
  dg_status s = dg_create(&md, &ld, &opt, &ctx_);
  if (s != DG_OK) {
    std::string msg = ctx_ ? dg_last_error(ctx_) : "dg_create failed";
    if (ctx_) dg_destroy(ctx_);
    ctx_ = nullptr;
    throw DGError(s, msg);
  }
  // integrator by order (DGAnalysis.cc:239-242)
  if (order == 0)
    integrator_ = new ForwardEuler(this);
  else if (order < 2)
    integrator_ = new RungeKuttaTVD2(this);
  else if (order < 3)
    integrator_ = new RungeKuttaTVD3(this);
  else
    integrator_ = new RungeKutta4(this);
}

DGAnalysis::~DGAnalysis() {
  delete integrator_;
  if (ctx_) dg_destroy(ctx_);
}

void DGAnalysis::check(dg_status s) const {
  if (s != DG_OK) throw DGError(s, dg_last_error(ctx_));
}

void DGAnalysis::setState(const std::vector<double>& U) { check(dg_set_state(ctx_, U.data(), U.size())); }

std::vector<double> DGAnalysis::getState() const {
  std::vector<double> U(dg_ndof(ctx_));
  check(dg_get_state(ctx_, U.data(), U.size()));
  return U;
}

void DGAnalysis::project(int field, const double* par, double t) { check(dg_project(ctx_, field, par, t)); }

std::vector<double> DGAnalysis::residual(double t) {
  std::vector<double> R(dg_ndof(ctx_));
  check(dg_residual(ctx_, t, R.data()));
  return R;
}

double DGAnalysis::computeDt(double cfl) {
  double dt = 0;
  check(dg_compute_dt(ctx_, cfl, &dt));
  return dt;
}

int DGAnalysis::run() {
  double dt = computeDt(CFLMAX);
  int N = (int)std::ceil((TEND - TACT) / dt);
  dt = (TEND - TACT) / N;
  check(dg_advance(ctx_, integrator_->scheme(), TACT, dt, N, nullptr));
  TACT = TEND;
  return N;
}

}  // namespace dgpar
