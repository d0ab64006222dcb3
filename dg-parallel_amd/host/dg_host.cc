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

ScalarLaw::ScalarLaw(FieldEvaluator* far, FieldEvaluator* velocity, FieldEvaluator* exact) {
  farField = far;
  exactField = exact;
  velField_ = velocity;
  for (int i = 0; i < 3; i++) vel_[i] = 0.0;
}

// the velocity at a point: the field's at time 0 (ScalarLaw.cc:33, 73), or the constant
void ScalarLaw::velocity(const mPoint& p, double* v) const {
  if (velField_) {
    v[0] = v[1] = v[2] = 0.0;
    velField_->eval(p, 0.0, v);
  } else {
    for (int d = 0; d < 3; d++) v[d] = vel_[d];
  }
}

void ScalarLaw::Fi(const mPoint& p, double* Q, mVector* flux) const {
  double v[3];
  velocity(p, v);
  for (int d = 0; d < 3; d++) flux[0](d) = v[d] * Q[0];
}

void ScalarLaw::riemannSolver(mVector& n, const mPoint& p, double* uL, double* uR, double* flux) const {
  double v[3];
  velocity(p, v);
  dgp::ScalarFieldLawD::riemann_at<double>(v, n.v, uL, uR, flux);
}

void ScalarLaw::boundaryFlux(mVector& n, int, const mPoint& p, double* Q, double* flux, double) const {
  double v[3], g[1];
  velocity(p, v);
  dgp::LawParams L = params_of(describe());
  dgp::ScalarLawD::ghost<double>(L, dgp::BC_FARFIELD, n.v, Q, g);
  dgp::ScalarFieldLawD::riemann_at<double>(v, n.v, Q, g, flux);
}

double ScalarLaw::maximumEigenValue(const double*, const mPoint& p) const {
  double v[3];
  velocity(p, v);
  return std::fabs(v[0]) + std::fabs(v[1]) + std::fabs(v[2]);
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
template <class LAW, int NF>
static void law_fi(const dg_law_desc& d, double* Q, mVector* flux) {
  dgp::LawParams L = params_of(d);
  double F[NF][2];
  LAW::template flux<double>(L, Q, F);
  for (int f = 0; f < NF; f++) {
    flux[f](0) = F[f][0];
    flux[f](1) = F[f][1];
    flux[f](2) = 0.0;
  }
}

static int bc_code(int BoundaryId) {
  return BoundaryId == DG_TAG_WALL || BoundaryId == DG_TAG_CURVED_WALL ? dgp::BC_WALL
         : BoundaryId == DG_TAG_FARFIELD2                             ? dgp::BC_FARFIELD2
                                                                     : dgp::BC_FARFIELD;
}

void Maxwell2d::Fi(const mPoint&, double* Q, mVector* flux) const { law_fi<dgp::Maxwell2dLawD, 3>(describe(), Q, flux); }
void Maxwell2d::riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const {
  dgp::Maxwell2dLawD::riemann<double>(params_of(describe()), n.v, uL, uR, flux);
}
void Maxwell2d::boundaryFlux(mVector& n, int BoundaryId, const mPoint&, double* Q, double* flux, double) const {
  dgp::Maxwell2dLawD::boundary<double>(params_of(describe()), bc_code(BoundaryId), n.v, Q, flux);
}
dg_law_desc Maxwell2d::describe() const {
  dg_law_desc d{};
  d.law = DG_LAW_MAXWELL2D;
  d.flux = DG_FLUX_UPWIND;
  far_state(farField, 3, d.farfield);
  return d;
}

void Burgers2D::Fi(const mPoint&, double* Q, mVector* flux) const { law_fi<dgp::Burgers2DLawD, 1>(describe(), Q, flux); }
void Burgers2D::riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* flux) const {
  dgp::Burgers2DLawD::riemann<double>(params_of(describe()), n.v, uL, uR, flux);
}
void Burgers2D::boundaryFlux(mVector& n, int BoundaryId, const mPoint&, double* Q, double* flux, double) const {
  dgp::Burgers2DLawD::boundary<double>(params_of(describe()), bc_code(BoundaryId), n.v, Q, flux);
}
dg_law_desc Burgers2D::describe() const {
  dg_law_desc d{};
  d.law = DG_LAW_BURGERS2D;
  d.flux = DG_FLUX_CENTRAL;
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
  dgp::LawParams L = params_of(describe());
  if (BoundaryId == DG_TAG_CHARACTERISTIC) {
    // the physical flux of the characteristic boundary state (EulerLaw.cc:185-278)
    dgp::EulerCharLawD::boundary<double>(L, dgp::BC_CHAR, n.v, Q, flux);
    return;
  }
  const int code = BoundaryId == DG_TAG_WALL || BoundaryId == DG_TAG_CURVED_WALL ? dgp::BC_WALL
                   : BoundaryId == DG_TAG_FARFIELD2                                ? dgp::BC_FARFIELD2
                                                                                  : dgp::BC_FARFIELD;
  dgp::EulerLawD::boundary<double>(L, code, n.v, Q, flux);
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

Mesh Mesh::structuredTet(int nx, int ny, int nz, double lx, double ly, double lz) {
  Mesh m;
  m.dim = 3;
  size_t E = (size_t)6 * nx * ny * nz;
  m.vids.resize(E * 4);
  m.xyz.resize(E * 12);
  dg_status s = dg_mesh_tet_structured(nx, ny, nz, lx, ly, lz, m.vids.data(), m.xyz.data());
  if (s != DG_OK) throw DGError(s, "dg_mesh_tet_structured rejected its arguments");
  return m;
}

Mesh Mesh::readGmsh(const char* path, bool split_quads) {
  dg_gmsh_info info{};
  char err[512] = {0};
  dg_status s = dg_mesh_read_gmsh(path, split_quads ? 1 : 0, &info, nullptr, nullptr, nullptr, err, sizeof(err));
  if (s != DG_OK) throw DGError(s, err);
  Mesh m;
  m.dim = info.dim;
  const size_t nv = (size_t)info.dim + 1;
  m.vids.resize(info.n_elems * nv);
  m.bc.resize(info.n_elems * nv);
  m.xyz.resize(info.n_elems * nv * info.dim);
  s = dg_mesh_read_gmsh(path, split_quads ? 1 : 0, &info, m.vids.data(), m.xyz.data(), m.bc.data(), err, sizeof(err));
  if (s != DG_OK) throw DGError(s, err);
  return m;
}

dg_mesh_desc Mesh::describe() const {
  dg_mesh_desc md{};
  md.dim = dim;
  md.n_elems = n_elems();
  md.vids = vids.data();
  md.xyz = xyz.data();
  md.bc = bc.empty() ? nullptr : bc.data();
  md.order = order.empty() ? nullptr : order.data();
  return md;
}

void Mesh::writeGmsh(const char* path) const {
  const dg_mesh_desc md = describe();
  char err[512] = {0};
  dg_status s = dg_mesh_write_gmsh(&md, path, err, sizeof(err));
  if (s != DG_OK) throw DGError(s, err);
}

// ---------------------------------------------------------------------------
dg_law_desc Euler2dTwoStates::describe() const {
  dg_law_desc d = Euler2dRoe::describe();
  for (int i = 0; i < 4; i++) d.farfield[4 + i] = second_[i];
  return d;
}

void Maxwell3d::Fi(const mPoint&, double* Q, mVector* F) const {
  const double *E = Q, *H = Q + 3;
  F[0] = mVector(0, -H[2], H[1]);
  F[1] = mVector(H[2], 0, -H[0]);
  F[2] = mVector(-H[1], H[0], 0);
  F[3] = mVector(0, E[2], -E[1]);
  F[4] = mVector(-E[2], 0, E[0]);
  F[5] = mVector(E[1], -E[0], 0);
}

void Maxwell3d::riemannSolver(mVector& n, const mPoint&, double* uL, double* uR, double* out) const {
  double Ea[3], Ha[3], dE[3], dH[3];
  for (int i = 0; i < 3; i++) {
    Ea[i] = 0.5 * (uL[i] + uR[i]);
    Ha[i] = 0.5 * (uL[3 + i] + uR[3 + i]);
    dE[i] = uR[i] - uL[i];
    dH[i] = uR[3 + i] - uL[3 + i];
  }
  out[0] = -(n(1) * Ha[2] - n(2) * Ha[1]);
  out[1] = -(n(2) * Ha[0] - n(0) * Ha[2]);
  out[2] = -(n(0) * Ha[1] - n(1) * Ha[0]);
  out[3] = n(1) * Ea[2] - n(2) * Ea[1];
  out[4] = n(2) * Ea[0] - n(0) * Ea[2];
  out[5] = n(0) * Ea[1] - n(1) * Ea[0];
  if (flux_kind == DG_FLUX_UPWIND) {
    const double ndE = n(0) * dE[0] + n(1) * dE[1] + n(2) * dE[2];
    const double ndH = n(0) * dH[0] + n(1) * dH[1] + n(2) * dH[2];
    for (int i = 0; i < 3; i++) {
      out[i] += 0.5 * (n(i) * ndE - dE[i]);
      out[3 + i] += 0.5 * (n(i) * ndH - dH[i]);
    }
  }
}

void Maxwell3d::boundaryFlux(mVector& n, int, const mPoint& x, double* Q, double* flux, double) const {
  double g[6];
  for (int i = 0; i < 3; i++) {
    g[i] = -Q[i];
    g[3 + i] = Q[3 + i];
  }
  riemannSolver(n, x, Q, g, flux);
}

dg_law_desc Maxwell3d::describe() const {
  dg_law_desc d{};
  d.law = DG_LAW_MAXWELL3D;
  d.flux = flux_kind;
  return d;
}

std::vector<uint8_t> DGAnalysis::limit() {
  std::vector<uint8_t> f(n_elems_);
  check(dg_limit(ctx_, f.data()));
  return f;
}

std::vector<int32_t> DGAnalysis::orders() const {
  std::vector<int32_t> o(n_elems_);
  check(dg_get_orders(ctx_, o.data()));
  return o;
}

std::vector<int32_t> DGAnalysis::adaptP(const dg_sensor_opts& opts) {
  std::vector<int32_t> o(n_elems_);
  check(dg_select_orders(ctx_, &opts, 1, o.data()));
  return o;
}

dg_tune_report DGAnalysis::autotune(double dt, int steps, const char* cache, bool force) {
  dg_tune_report r{};
  check(dg_autotune(ctx_, integrator_->scheme(), dt, steps, cache, force ? DG_TUNE_FORCE : 0, &r));
  return r;
}

dg_schedule DGAnalysis::schedule() const {
  dg_schedule s{};
  check(dg_get_schedule(ctx_, &s));
  return s;
}

// ---------------------------------------------------------------------------
double TimeIntegrator::advanceInTime(double t, double dt) {
  double r = 0.0;
  an_->check(dg_advance(an_->context(), scheme_, t, dt, 1, &r));
  return r;
}

// a FieldEvaluator behind the C ABI's dg_field_fn
static void eval_field(const double* x, double t, double* val, void* user) {
  static_cast<const FieldEvaluator*>(user)->eval(mPoint(x[0], x[1], x[2]), t, val);
}

DGAnalysis::DGAnalysis(const Mesh& mesh, ConservationLaw* law, DGLimiter* limiter, int order, int precision)
    : mesh_(mesh), law_(law), order_(order) {
  const dg_mesh_desc md = mesh.describe();
  n_elems_ = md.n_elems;
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
  if (limiter) check(dg_set_limiter(ctx_, 1, limiter->tvbM(), limiter->nu()));
  if (FieldEvaluator* vf = law->velocityField()) check(dg_set_velocity_field(ctx_, eval_field, vf));
  // every cell starts as the L2 projection of the law's far field (DGAnalysis.cc:200)
  if (FieldEvaluator* ff = law->getFarField()) check(dg_project_fn(ctx_, eval_field, ff, 0.0));
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

int DGAnalysis::adaptH(const std::vector<int8_t>& marks, int max_level) {
  if ((int)marks.size() != n_elems_) throw DGError(DG_ERR_ARG, "adaptH: one mark per cell");
  int32_t n = 0;
  check(dg_adapt_h(ctx_, marks.data(), max_level, &n));
  n_elems_ = n;
  return n;
}

Mesh DGAnalysis::refinedMesh() const {
  Mesh m;
  m.dim = 2;
  int32_t n = 0;
  check(dg_get_mesh(ctx_, &n, nullptr, nullptr, nullptr, nullptr));
  m.vids.resize((size_t)n * 3);
  m.xyz.resize((size_t)n * 6);
  m.bc.resize((size_t)n * 3);
  check(dg_get_mesh(ctx_, nullptr, m.vids.data(), m.xyz.data(), m.bc.data(), nullptr));
  return m;
}

std::vector<int32_t> DGAnalysis::sortCellsBySize(int max_levels) {
  std::vector<int32_t> lev(n_elems_);
  check(dg_lts_setup(ctx_, max_levels, lev.data()));
  delete integrator_;
  integrator_ = new RungeKuttaTVD2adaptive(this);
  return lev;
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

void DGAnalysis::write(const char* mesh_path, const char* data_path) const {
  mesh_.writeGmsh(mesh_path);
  check(dg_write_restart(ctx_, data_path, TACT, ITER));
}

void DGAnalysis::read(const char* data_path) { check(dg_read_restart(ctx_, data_path, &TACT, &ITER)); }

void DGAnalysis::exportGmsh(const char* fileName, int iteration) const {
  check(dg_export_pos(ctx_, fileName, iteration, 0, TACT, -1, nullptr));
}

int DGAnalysis::run() {
  double dt = computeDt(CFLMAX);
  int N = (int)std::ceil((TEND - TACT) / dt);
  dt = (TEND - TACT) / N;
  check(dg_advance(ctx_, integrator_->scheme(), TACT, dt, N, nullptr));
  TACT = TEND;
  ITER += N;
  return N;
}

}  // namespace dgpar
