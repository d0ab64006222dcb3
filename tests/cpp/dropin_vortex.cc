// Drives the library through the host C++ mirror classes the way a reference
// scenario drives DG-Parallel (Examples.cc / Lilia.cc): build a law and a mesh,
// construct DGAnalysis, then call residual and TimeIntegrator::advanceInTime.
//   dropin_vortex OUT_PREFIX [vortex|tube|maxwell|restart|scalarpb]
// writes OUT_PREFIX.U0 (projected state), .R (residual), .U2 (after two steps of
// the integrator DGAnalysis picked for the order), raw little-endian doubles, and
// prints dt and the residual norms.
//   vortex   Euler2d (LF), P3, periodic jittered 10x10x2 mesh
//   tube     Euler2dTwoStates (Roe + entropy fix) with DGTriangleLimiter, seeded
//            orders P1-P4 on a 16x4x2 tube, the state limited once, two SSP-RK3
//            steps; then adaptP and .O (new orders as doubles), .U3 (state after it)
//   maxwell  Maxwell3d (upwind), P3 on a 2x2x2 Kuhn mesh, cavity mode, RK4
//   scalarpb the reference's ScalarPb: ScalarLaw with the rotating velocity
//            FieldEvaluator, P1 on a periodic jittered 12x12x2 mesh of [-1,1]^2
//   restart  the vortex written, plotted, read back and continued (see restart())
//   lts      the vortex (Euler2dLF, P1) on a graded periodic mesh: sortCellsBySize(3)
//            (integrator RungeKuttaTVD2adaptive), two steps of the coarsest level's dt;
//            .L (levels as doubles), .U2
//   adapt    the vortex (Euler2dLF, P2): adaptH splits the cells within 2 of the
//            centre, two RK3 steps on the refined mesh; .M (refined corners
//            [E][3][2]), .U2
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../dg-parallel_amd/host/dg_host.h"

static void dump(const std::string& path, const std::vector<double>& v) {
  FILE* f = fopen(path.c_str(), "wb");
  fwrite(v.data(), sizeof(double), v.size(), f);
  fclose(f);
}

static int vortex(const std::string& out) {
  const int32_t tags[4] = {0, 0, 0, 0};
  dgpar::Mesh mesh = dgpar::Mesh::structuredTri(10, 10, 0.0, 0.0, 10.0, 10.0, true, true, 0.1, 77, tags);
  dgpar::ConstantField far({1.0, 1.0, 1.0, 3.5});
  dgpar::Euler2d law(&far, 1.4);
  dgpar::DGAnalysis an(mesh, &law, nullptr, 3);
  const double vortex[6] = {5.0, 5.0, 5.0, 1.0, 1.0, 10.0};
  an.project(DG_FIELD_VORTEX, vortex);
  dump(out + ".U0", an.getState());
  dump(out + ".R", an.residual(0.0));
  double dt = an.computeDt(0.1);
  double r1 = an.integrator()->advanceInTime(0.0, dt);
  double r2 = an.integrator()->advanceInTime(dt, dt);
  dump(out + ".U2", an.getState());
  printf("dt %.17e\nresid %.17e %.17e\n", dt, r1, r2);
  return 0;
}

static int tube(const std::string& out) {
  const int32_t tags[4] = {DG_TAG_WALL, DG_TAG_FARFIELD2, DG_TAG_WALL, DG_TAG_FARFIELD};
  dgpar::Mesh mesh = dgpar::Mesh::structuredTri(16, 4, 0.0, 0.0, 1.0, 0.25, false, false, 0.1, 5, tags);
  // seeded orders 1..4 (splitmix64 of the element index)
  mesh.order.resize(mesh.n_elems());
  for (int e = 0; e < mesh.n_elems(); e++) {
    uint64_t z = (uint64_t)e * 0x9E3779B97F4A7C15ull + 11;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    mesh.order[e] = 1 + (int)((z ^ (z >> 31)) % 4);
  }
  dgpar::ConstantField far({1.0, 0.0, 0.0, 2.5});
  const double right[4] = {0.125, 0.0, 0.0, 0.25};
  dgpar::Euler2dTwoStates law(&far, right, 1.4, true);
  dgpar::DGTriangleLimiter lim(0.0, 1.0);
  dgpar::DGAnalysis an(mesh, &law, &lim, 2);  // order 2: the reference picks SSP-RK3
  const double sod[1] = {0.53};
  an.project(DG_FIELD_SOD, sod);
  an.limit();
  dump(out + ".U0", an.getState());
  dump(out + ".R", an.residual(0.0));
  double dt = an.computeDt(0.1);
  double r1 = an.integrator()->advanceInTime(0.0, dt);
  double r2 = an.integrator()->advanceInTime(dt, dt);
  dump(out + ".U2", an.getState());
  dg_sensor_opts so{{0.05, 0.02, 0.01, 0.005}, {5e-5, 2e-5, 1e-5, 5e-6}, 1, 4};
  std::vector<int32_t> o = an.adaptP(so);
  dump(out + ".O", std::vector<double>(o.begin(), o.end()));
  dump(out + ".U3", an.getState());
  std::vector<int32_t> o0 = mesh.order;
  dump(out + ".O0", std::vector<double>(o0.begin(), o0.end()));
  printf("dt %.17e\nresid %.17e %.17e\n", dt, r1, r2);
  return 0;
}

// the reference's ScalarPb scenario (Lilia.cc:84-108): ScalarLaw with the
// rotating VelocityOne field (Lilia.cc:66-82) as its velocity FieldEvaluator, on
// a jittered periodic 12x12x2 mesh of [-1,1]^2, P1 (the reference's benchmark
// order), the integrator DGAnalysis picks for it
class VelocityOne : public dgpar::FieldEvaluator {
 public:
  void eval(const dgpar::mPoint& space, double, double* val) const override {
    const double pi = 3.14159265358979323846;
    val[0] = 2. * pi * space(1);
    val[1] = -2. * pi * space(0);
    val[2] = 0.0;
  }
  int order() const override { return 1; }
};

// its initial condition (CaseOne, Lilia.cc:38-58): the far field the
// constructor projects into every cell
class CaseOne : public dgpar::FieldEvaluator {
 public:
  void eval(const dgpar::mPoint& space, double, double* val) const override {
    const double pi = 3.14159265358979323846;
    const double r = std::sqrt((space(0) + 0.5) * (space(0) + 0.5) + space(1) * space(1));
    val[0] = r <= 0.25 ? std::pow(std::cos(2 * pi * r), 2) : 0.0;
  }
  int order() const override { return 2; }
};

static int scalarpb(const std::string& out) {
  const int32_t tags[4] = {0, 0, 0, 0};
  dgpar::Mesh mesh = dgpar::Mesh::structuredTri(12, 12, -1.0, -1.0, 2.0, 2.0, true, true, 0.1, 31, tags);
  CaseOne i;
  VelocityOne v;
  dgpar::ScalarLaw theLaw(&i, &v, &i);
  dgpar::DGAnalysis analysis(mesh, &theLaw, nullptr, 1);
  dump(out + ".UP", analysis.getState());  // the constructor's projection of CaseOne
  std::vector<double> U(analysis.ndof());
  for (size_t k = 0; k < U.size(); k++) U[k] = std::sin(0.37 * (double)k) + 0.2 * std::cos(1.3 * (double)k);
  analysis.setState(U);
  dump(out + ".U0", analysis.getState());
  dump(out + ".R", analysis.residual(0.0));
  double dt = analysis.computeDt(0.1);
  double r1 = analysis.integrator()->advanceInTime(0.0, dt);
  double r2 = analysis.integrator()->advanceInTime(dt, dt);
  dump(out + ".U2", analysis.getState());
  printf("dt %.17e\nresid %.17e %.17e\n", dt, r1, r2);
  return 0;
}

static int maxwell(const std::string& out) {
  dgpar::Mesh mesh = dgpar::Mesh::structuredTet(2, 2, 2, 1.0, 1.0, 1.0);
  dgpar::Maxwell3d law(true);
  dgpar::DGAnalysis an(mesh, &law, nullptr, 3);
  an.project(DG_FIELD_CAVITY, nullptr);
  dump(out + ".U0", an.getState());
  dump(out + ".R", an.residual(0.0));
  double dt = an.computeDt(0.1);
  double r1 = an.integrator()->advanceInTime(0.0, dt);
  double r2 = an.integrator()->advanceInTime(dt, dt);
  dump(out + ".U2", an.getState());
  printf("dt %.17e\nresid %.17e %.17e\n", dt, r1, r2);
  return 0;
}

// checkpoint, plot and restart through the mirror (DGAnalysis::write / read /
// exportGmsh): after two steps the run writes OUT_PREFIX.msh / .dat and the .pos
// views, a second DGAnalysis built from the written mesh file reads the data,
// and both take two more steps: .U4a (first) and .U4b (restarted)
static int restart(const std::string& out) {
  const int32_t tags[4] = {0, 0, 0, 0};
  dgpar::Mesh mesh = dgpar::Mesh::structuredTri(10, 10, 0.0, 0.0, 10.0, 10.0, true, true, 0.1, 77, tags);
  dgpar::ConstantField far({1.0, 1.0, 1.0, 3.5});
  dgpar::Euler2d law(&far, 1.4);
  dgpar::DGAnalysis an(mesh, &law, nullptr, 3);
  const double vortex[6] = {5.0, 5.0, 5.0, 1.0, 1.0, 10.0};
  an.project(DG_FIELD_VORTEX, vortex);
  double dt = an.computeDt(0.1);
  an.integrator()->advanceInTime(0.0, dt);
  an.integrator()->advanceInTime(dt, dt);
  an.TACT = 2 * dt;
  an.ITER = 2;
  an.write((out + ".msh").c_str(), (out + ".dat").c_str());
  an.exportGmsh(out.c_str(), an.ITER);
  dgpar::DGAnalysis again(dgpar::Mesh::readGmsh((out + ".msh").c_str()), &law, nullptr, 3);
  again.read((out + ".dat").c_str());
  printf("dt %.17e\nrestart %.17e %d\n", dt, again.TACT, again.ITER);
  for (int n = 2; n < 4; n++) {
    an.integrator()->advanceInTime(n * dt, dt);
    again.integrator()->advanceInTime(n * dt, dt);
  }
  dump(out + ".U4a", an.getState());
  dump(out + ".U4b", again.getState());
  return 0;
}

// local time stepping: a tensor grid of [0,10]^2 with cell widths 4:2:1:1:2:4 in
// both directions (blocks of 2 cells), P1 vortex, SSP-RK2 with local steps
static int lts(const std::string& out) {
  const int n = 12;
  std::vector<double> xs(n + 1, 0.0);
  const double w[6] = {4, 2, 1, 1, 2, 4};
  for (int i = 0; i < n; i++) xs[i + 1] = xs[i] + w[i / 2];
  for (double& x : xs) x *= 10.0 / 28.0;
  dgpar::Mesh mesh;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      const int c[2][3][2] = {{{0, 0}, {1, 0}, {1, 1}}, {{0, 0}, {1, 1}, {0, 1}}};
      for (int t = 0; t < 2; t++)
        for (int v = 0; v < 3; v++) {
          const int a = i + c[t][v][0], b = j + c[t][v][1];
          mesh.vids.push_back((b % n) * (n + 1) + (a % n));
          mesh.xyz.push_back(xs[a]);
          mesh.xyz.push_back(xs[b]);
        }
    }
  dgpar::ConstantField far({1.0, 1.0, 1.0, 3.5});
  dgpar::Euler2dLF law(&far, 1.4);
  dgpar::DGAnalysis an(mesh, &law, nullptr, 1);
  const double vortex[6] = {5.0, 5.0, 5.0, 1.0, 1.0, 10.0};
  an.project(DG_FIELD_VORTEX, vortex);
  const std::vector<int32_t> lev = an.sortCellsBySize(3);
  dump(out + ".U0", an.getState());
  dump(out + ".L", std::vector<double>(lev.begin(), lev.end()));
  const double dt = an.computeDt(0.1);
  const double r1 = an.integrator()->advanceInTime(0.0, dt);
  const double r2 = an.integrator()->advanceInTime(dt, dt);
  dump(out + ".U2", an.getState());
  printf("dt %.17e\nresid %.17e %.17e\n", dt, r1, r2);
  return 0;
}

// h-adaptivity: the P2 vortex on the periodic jittered 10x10x2 mesh of [0,10]^2,
// the cells whose centroid lies within 2 of the centre split (adaptH, at most one
// level), then two SSP-RK3 steps on the refined mesh
static int adapt(const std::string& out) {
  const int32_t tags[4] = {0, 0, 0, 0};
  dgpar::Mesh mesh = dgpar::Mesh::structuredTri(10, 10, 0.0, 0.0, 10.0, 10.0, true, true, 0.1, 77, tags);
  dgpar::ConstantField far({1.0, 1.0, 1.0, 3.5});
  dgpar::Euler2dLF law(&far, 1.4);
  dgpar::DGAnalysis an(mesh, &law, nullptr, 2);
  const double vortex[6] = {5.0, 5.0, 5.0, 1.0, 1.0, 10.0};
  an.project(DG_FIELD_VORTEX, vortex);
  std::vector<int8_t> marks(mesh.n_elems(), 0);
  for (int e = 0; e < mesh.n_elems(); e++) {
    double cx = 0, cy = 0;
    for (int v = 0; v < 3; v++) {
      cx += mesh.xyz[e * 6 + v * 2] / 3.0;
      cy += mesh.xyz[e * 6 + v * 2 + 1] / 3.0;
    }
    marks[e] = std::hypot(cx - 5.0, cy - 5.0) < 2.0 ? 1 : 0;
  }
  an.adaptH(marks, 1);
  const dgpar::Mesh fine = an.refinedMesh();
  dump(out + ".M", fine.xyz);
  dump(out + ".U0", an.getState());
  const double dt = an.computeDt(0.1);
  const double r1 = an.integrator()->advanceInTime(0.0, dt);
  const double r2 = an.integrator()->advanceInTime(dt, dt);
  dump(out + ".U2", an.getState());
  printf("dt %.17e\nresid %.17e %.17e\n", dt, r1, r2);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::string out = argv[1], mode = argc > 2 ? argv[2] : "vortex";
  try {
    if (mode == "tube") return tube(out);
    if (mode == "maxwell") return maxwell(out);
    if (mode == "restart") return restart(out);
    if (mode == "scalarpb") return scalarpb(out);
    if (mode == "lts") return lts(out);
    if (mode == "adapt") return adapt(out);
    return vortex(out);
  } catch (const dgpar::DGError& e) {
    fprintf(stderr, "DGError %d: %s\n", (int)e.status, e.what());
    return 1;
  }
}
