// Drives the library through the host C++ mirror classes the way a reference
// scenario drives DG-Parallel (Examples.cc / Lilia.cc): build a law and a mesh,
// construct DGAnalysis, then call residual and TimeIntegrator::advanceInTime.
//   dropin_vortex OUT_PREFIX
// writes OUT_PREFIX.U0 (projected state), .R (residual), .U2 (after two steps of
// the integrator DGAnalysis picked for P3, RK4), raw little-endian doubles, and
// prints dt and the residual norms.
#include <cstdio>
#include <string>
#include <vector>

#include "../../dg-parallel_amd/host/dg_host.h"

static void dump(const std::string& path, const std::vector<double>& v) {
  FILE* f = fopen(path.c_str(), "wb");
  fwrite(v.data(), sizeof(double), v.size(), f);
  fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::string out = argv[1];
  try {
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
  } catch (const dgpar::DGError& e) {
    fprintf(stderr, "DGError %d: %s\n", (int)e.status, e.what());
    return 1;
  }
  return 0;
}
