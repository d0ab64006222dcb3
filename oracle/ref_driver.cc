// ref_driver.cc -- a serial driver for the reference's own DG classes, compiled
// against the reference sources (oracle/Makefile.ref) into oracle/_ref/.  TEST
// INFRASTRUCTURE ONLY: it produces reference residuals / RK steps so the oracle
// can be checked against the real reference code.  It plays the role of the
// reference's main.cc (main.cc:55-159) and scenario functions (Lilia.cc:84-108):
//
//   ref_driver MESH.msh LAW FLUX ORDER OP STATE_IN OUT [DT NSTEPS]
//     LAW   scalar | euler | maxwell2d | burgers
//     FLUX  upwind | exact | roe | lf (Euler; the other laws have one flux each)
//     OP    residual | fe | tvd2 | tvd3 | rk4 | write (DGAnalysis::write of the
//           state: the data stream to OUT, the mesh to OUT.msh) | project (the
//           coefficients the constructor projected from the far field)
//     STATE_IN: one line per cell: 3 vertex ids (any order), then Nf*Np coefficients;
//          with ORDER 0 each line carries the cell's order after the ids, and the
//          driver gives every cell that order through DGCell::adaptOrder
//          (DGCell.cc:1542-1602), which re-initialises the adjacent faces with the
//          rule of degree 2 max(pL, pR) + 1 (DGCell.cc:408-423)
//     OUT: one line per cell in the reference's iteration order: its 3 vertex ids
//          in the reference's own vertex order, then the residual or new state
//   scalar velocity (VX VY) and far-field state are read from the environment
//   (REF_VEL="vx vy", or REF_VEL=rotate for ScalarPb's (2 pi y, -2 pi x);
//   REF_FARFIELD="q0 q1 q2 q3"); time steps start at REF_T0
//   (default 0; Maxwell2d's RHS holds a pulse source for t <= 0.06).
//   REF_FAR=bump: the scalar law's far field is ScalarPb's initial bump.
//   REF_SWEEPS=N (residual): N timed sweeps before the written residual,
//   "residual_seconds S N" on stdout.
//
// Only the serial code paths run (numThreads = 1): the threaded face assembly
// races (parallel.cc:46-76), the serial one does not.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <string>
#include <set>
#include <vector>

#include "ConservationLaw.h"
#include "DGAnalysis.h"
#include "DGCell.h"
#include "EulerLaw.h"
#include "Burgers.h"
#include "FieldEvaluator.h"
#include "MaxwellLaw.h"
#include "ScalarLaw.h"
#include "TimeIntegrator.h"
#include "Vertex.h"
#include "mDGMesh.h"
#include "mEntity.h"
#include "mImportExport.h"
#include "mPoint.h"
#include "mVector.h"
#include "parallel.h"

// globals the reference's main.cc defines (main.cc:32-36)
unsigned int numCPU = 1;
unsigned int numThreads = 1;
vector<infoWrapper> threadInfo[3];
double timeRK = 0;
timespec timer1load, timer2load;
timespec timer_start, timer_stop;

static double g_vel[3] = {1.0, 0.5, 0.0};
static double g_ff[8] = {0, 0, 0, 0, 0, 0, 0, 0};

class ConstField : public FieldEvaluator {
  const double* v;
  int n;

 public:
  ConstField(const double* vals, int nv) : v(vals), n(nv) {}
  virtual void eval(const mPoint&, double, double* val) const {
    for (int i = 0; i < n; i++) val[i] = v[i];
  }
  virtual int order() const { return 0; }
};

// the rotating velocity of the reference's ScalarPb scenario (VelocityOne,
// Lilia.cc:66-82): (2 pi y, -2 pi x), selected by REF_VEL=rotate
class RotatingField : public FieldEvaluator {
 public:
  virtual void eval(const mPoint& x, double, double* val) const {
    const double pi = 3.14159265358979323846;
    val[0] = 2. * pi * x(1);
    val[1] = -2. * pi * x(0);
    val[2] = 0.0;
  }
  virtual int order() const { return 1; }
};

// the initial condition of the reference's ScalarPb scenario (CaseOne,
// Lilia.cc:38-58): cos^2(2 pi r) within r <= 0.25 of (-0.5, 0), zero outside;
// the scalar law's far field with REF_FAR=bump
class BumpField : public FieldEvaluator {
 public:
  virtual void eval(const mPoint& x, double, double* val) const {
    const double pi = 3.14159265358979323846;
    const double r = sqrt((x(0) + 0.5) * (x(0) + 0.5) + x(1) * x(1));
    val[0] = r <= 0.25 ? cos(2 * pi * r) * cos(2 * pi * r) : 0.0;
  }
  virtual int order() const { return 2; }
};

// Local Lax-Friedrichs flux plugged into the reference's Euler2d law through
// the riemannSolver virtual (ConservationLaw.h:29-33): the only part of the
// path the reference does not provide.  Wave speed as maximumEigenValue(n,.,.)
// (EulerLaw.cc:1022-1034).
class Euler2dLF : public Euler2d {
 public:
  Euler2dLF(set<int>& w, FieldEvaluator* f, double g) : Euler2d(w, f, g) {}
  virtual void riemannSolver(mVector& n, const mPoint& p, double* uL, double* uR, double* flux) const {
    mVector fL[4], fR[4];
    Fi(p, uL, fL);
    Fi(p, uR, fR);
    double lam = maximumEigenValue(n, uL, uR);
    for (int i = 0; i < 4; i++) flux[i] = 0.5 * ((fL[i] + fR[i]) * n - lam * (uR[i] - uL[i]));
  }
};

static int np_tri(int p) { return (p + 1) * (p + 2) / 2; }

int main(int argc, char** argv) {
  if (argc < 8) {
    fprintf(stderr, "usage: %s MESH LAW FLUX ORDER OP STATE_IN OUT [DT NSTEPS]\n", argv[0]);
    return 2;
  }
  const char* law_s = argv[2];
  const char* flux_s = argv[3];
  int order = atoi(argv[4]);
  const char* op = argv[5];
  double dt = argc > 8 ? atof(argv[8]) : 0.0;
  int nsteps = argc > 9 ? atoi(argv[9]) : 1;
  const bool rotate = getenv("REF_VEL") && !strcmp(getenv("REF_VEL"), "rotate");
  if (getenv("REF_VEL") && !rotate) sscanf(getenv("REF_VEL"), "%lf %lf", &g_vel[0], &g_vel[1]);
  if (getenv("REF_FARFIELD"))
    sscanf(getenv("REF_FARFIELD"), "%lf %lf %lf %lf", &g_ff[0], &g_ff[1], &g_ff[2], &g_ff[3]);

  mImportExport io;
  mDGMesh* mesh = new mDGMesh;
  io.import(argv[1], mesh);
  const int dim = 2;
  mesh->createCellBoundaries(dim, dim - 1);
  mesh->createConnections(dim - 1, dim);
  mesh->createConnections(0, dim - 1);
  mesh->setPeriodicBC(dim);
  for (int i = 0; i < 3; i++) threadInfo[i].push_back(infoWrapper(mesh->begin(i), mesh->end(i), 0, 0));

  double t0 = getenv("REF_T0") ? atof(getenv("REF_T0")) : 0.0;
  ConservationLaw* law = 0;
  // far-field evaluators of each law's field count (the laws evaluate into
  // arrays of that size)
  ConstField vel(g_vel, 3), ff(g_ff, 4), ff3(g_ff, 3), ff1(g_ff, 1);
  set<int> walls;
  walls.insert(40000);
  int nf;
  RotatingField rot;
  BumpField bump;
  const bool bump_far = getenv("REF_FAR") && !strcmp(getenv("REF_FAR"), "bump");
  if (!strcmp(law_s, "scalar")) {
    FieldEvaluator* far = bump_far ? (FieldEvaluator*)&bump : (FieldEvaluator*)&ff;
    law = new ScalarLaw(far, rotate ? (FieldEvaluator*)&rot : (FieldEvaluator*)&vel, far);
    nf = 1;
  } else if (!strcmp(law_s, "maxwell2d")) {
    law = new Maxwell2d(walls, &ff3);
    nf = 3;
  } else if (!strcmp(law_s, "burgers")) {
    law = new Burgers2D(&ff1, &ff1);
    nf = 1;
  } else {
    if (!strcmp(flux_s, "roe"))
      law = new Euler2dRoe(walls, &ff, 1.4);
    else if (!strcmp(flux_s, "lf"))
      law = new Euler2dLF(walls, &ff, 1.4);
    else
      law = new Euler2d(walls, &ff, 1.4);
    nf = 4;
  }
  const bool mixed = order == 0;
  DGAnalysis* an = new DGAnalysis(mesh, law, 0, mixed ? 1 : order);
  (void)an;

  // read the state keyed by sorted vertex ids
  std::map<std::vector<int>, std::vector<double> > state;
  std::map<std::vector<int>, int> orders;
  FILE* fi = fopen(argv[6], "r");
  if (!fi) return 3;
  int v[3];
  while (fscanf(fi, "%d %d %d", &v[0], &v[1], &v[2]) == 3) {
    std::vector<int> key(v, v + 3);
    std::sort(key.begin(), key.end());
    int p = order;
    if (mixed && fscanf(fi, "%d", &p) != 1) return 4;
    orders[key] = p;
    std::vector<double> c(nf * np_tri(p));
    for (size_t i = 0; i < c.size(); i++)
      if (fscanf(fi, "%lf", &c[i]) != 1) return 4;
    state[key] = c;
  }
  fclose(fi);
  int DOF = 0;
  for (mMesh::iter it = mesh->begin(dim); it != mesh->end(dim); ++it) {
    mEntity* m = *it;
    std::vector<int> key;
    for (int k = 0; k < 3; k++) key.push_back(m->get(0, k)->getId());
    std::sort(key.begin(), key.end());
    if (!state.count(key)) {
      fprintf(stderr, "cell not in state file\n");
      return 5;
    }
    if (mixed) ((DGCell*)m->getCell())->adaptOrder(orders[key]);
  }
  // op project: the coefficients DGAnalysis's constructor projected from the far
  // field (DGAnalysis.cc:200) are the output; the state file only gives the order
  const bool project = !strcmp(op, "project");
  for (mMesh::iter it = mesh->begin(dim); it != mesh->end(dim) && !project; ++it) {
    mEntity* m = *it;
    DGCell* cell = (DGCell*)m->getCell();
    std::vector<int> key;
    for (int k = 0; k < 3; k++) key.push_back(m->get(0, k)->getId());
    std::sort(key.begin(), key.end());
    const std::vector<double>& c = state[key];
    const int NP = np_tri(orders[key]);
    for (int f = 0; f < nf; f++)
      for (int k = 0; k < NP; k++) cell->theFieldsCoefficients->get(f, k) = c[f * NP + k];
    cell->computeMean();
    cell->ZeroRHS();
    cell->setTimeStep(dt);
    DOF += nf * NP;
  }

  if (!strcmp(op, "write")) {
    // DGAnalysis::write (DGPref.cc:712-740): the mesh to OUT.msh, the data stream to OUT
    std::string mp = std::string(argv[7]) + ".msh";
    std::ofstream ms(mp.c_str()), ds(argv[7]);
    an->write(ms, ds);
    return ms.good() && ds.good() ? 0 : 6;
  }
  const bool residual = !strcmp(op, "residual");
  if (project) {
  } else if (residual) {
    // serial assembly as TimeIntegrator::assembleVolume/assembleBoundary do with one thread
    const int sweeps = getenv("REF_SWEEPS") ? atoi(getenv("REF_SWEEPS")) : 0;
    if (sweeps > 0) {
      // timed sweeps of both passes, as the reference's profiling harness runs them
      // (DGAnalysis.cc:321-477, serial); the right-hand sides accumulate and are
      // cleared before the residual that is written
      timespec c0, c1;
      clock_gettime(CLOCK_MONOTONIC, &c0);
      for (int k = 0; k < sweeps; k++) {
        for (mMesh::iter it = mesh->begin(dim); it != mesh->end(dim); ++it)
          ((DGCell*)(*it)->getCell())->computeVolumeContribution(0.0);
        for (mMesh::iter it = mesh->begin(dim - 1); it != mesh->end(dim - 1); ++it)
          ((DGBoundaryCell*)(*it)->getCell())->computeBoundaryContributions(0.0);
      }
      clock_gettime(CLOCK_MONOTONIC, &c1);
      printf("residual_seconds %.6f %d\n", (c1.tv_sec - c0.tv_sec) + 1e-9 * (c1.tv_nsec - c0.tv_nsec), sweeps);
      for (mMesh::iter it = mesh->begin(dim); it != mesh->end(dim); ++it) ((DGCell*)(*it)->getCell())->ZeroRHS();
    }
    for (mMesh::iter it = mesh->begin(dim); it != mesh->end(dim); ++it)
      ((DGCell*)(*it)->getCell())->computeVolumeContribution(0.0);
    for (mMesh::iter it = mesh->begin(dim - 1); it != mesh->end(dim - 1); ++it)
      ((DGBoundaryCell*)(*it)->getCell())->computeBoundaryContributions(0.0);
  } else {
    TimeIntegrator* ti;
    if (!strcmp(op, "rk4"))
      ti = new RungeKutta4(mesh, 0, nf, DOF);
    else if (!strcmp(op, "tvd3"))
      ti = new RungeKuttaTVD3(mesh, 0, nf, DOF);
    else if (!strcmp(op, "tvd2"))
      ti = new RungeKuttaTVD2(mesh, 0, nf, DOF);
    else
      ti = new ForwardEuler(mesh, 0, nf, DOF);
    double t = t0;
    timespec c0, c1;
    clock_gettime(CLOCK_MONOTONIC, &c0);
    for (int s = 0; s < nsteps; s++) {
      double r = ti->advanceInTime(t, dt);
      printf("resid %.17e\n", r);
      t += dt;
    }
    clock_gettime(CLOCK_MONOTONIC, &c1);
    // the step loop alone (mesh import and DGAnalysis construction excluded)
    printf("step_seconds %.6f\n", (c1.tv_sec - c0.tv_sec) + 1e-9 * (c1.tv_nsec - c0.tv_nsec));
  }

  FILE* fo = fopen(argv[7], "w");
  if (!fo) return 6;
  for (mMesh::iter it = mesh->begin(dim); it != mesh->end(dim); ++it) {
    mEntity* m = *it;
    DGCell* cell = (DGCell*)m->getCell();
    std::vector<int> key;
    for (int k = 0; k < 3; k++) key.push_back(m->get(0, k)->getId());
    std::sort(key.begin(), key.end());
    const int NP = np_tri(orders[key]);
    for (int k = 0; k < 3; k++) fprintf(fo, "%d ", m->get(0, k)->getId());
    for (int f = 0; f < nf; f++)
      for (int k = 0; k < NP; k++)
        fprintf(fo, " %.17e", residual ? cell->theRightHandSide[f * NP + k] : cell->theFieldsCoefficients->get(f, k));
    fprintf(fo, "\n");
  }
  fclose(fo);
  return 0;
}
