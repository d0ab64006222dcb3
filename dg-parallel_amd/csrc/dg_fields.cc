// Exact fields of the benchmark problems, their L2 projection onto the modal
// basis (DGCell::L2ProjInitial, DGCell.cc:1458-1483) and per-field L2 errors
// (DGCell::densityErrorL2, DGError.cc:805-831, rule of degree 2(p+1)).
// Host FP64; used for initial conditions and accuracy reports, outside any
// timed region.
#include "dg_fields.h"

#include <cmath>

namespace dgp {

static double wrap(double x, double L) {
  double r = std::fmod(x, L);
  if (r < 0) r += L;
  return r;
}

int exact_field(int field, const double* par, const LawParams& law, const double* x, double t, double* val) {
  const double pi = 3.14159265358979323846;
  switch (field) {
    case FIELD_GAUSSIAN: {
      // exp(-((x-.5)^2+(y-.5)^2)/0.01) carried by the constant velocity, unit square periodic
      double cx = wrap(x[0] - law.vel[0] * t, 1.0) - 0.5, cy = wrap(x[1] - law.vel[1] * t, 1.0) - 0.5;
      val[0] = std::exp(-(cx * cx + cy * cy) / 0.01);
      return 0;
    }
    case FIELD_VORTEX: {
      // isentropic vortex: par = beta, xc, yc, u0, v0, periodic box length
      double g = law.gamma, beta = par[0], L = par[5], u0 = par[3], v0 = par[4];
      double xs = wrap(x[0] - u0 * t - par[1] + 0.5 * L, L) - 0.5 * L;
      double ys = wrap(x[1] - v0 * t - par[2] + 0.5 * L, L) - 0.5 * L;
      double r2 = xs * xs + ys * ys;
      double T = 1.0 - (g - 1.0) * beta * beta * std::exp(1.0 - r2) / (8.0 * g * pi * pi);
      double rho = std::pow(T, 1.0 / (g - 1.0));
      double du = beta / (2.0 * pi) * std::exp(0.5 * (1.0 - r2));
      double u = u0 - du * ys, v = v0 + du * xs;
      double p = std::pow(rho, g);
      val[0] = rho;
      val[1] = rho * u;
      val[2] = rho * v;
      val[3] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v);
      return 0;
    }
    case FIELD_SOD: {
      // ShockTube2 initial state (Examples.cc:62-83), interface x = par[0]
      bool left = x[0] < par[0];
      val[0] = left ? 1.0 : 0.125;
      val[1] = 0.0;
      val[2] = 0.0;
      val[3] = (left ? 1.0 : 0.1) / (law.gamma - 1.0);
      return 0;
    }
    case FIELD_CAVITY: {
      // PEC cavity mode (1,1,1) in the unit cube, E amplitude (1,-1,0), omega = pi sqrt 3
      double w = pi * std::sqrt(3.0);
      double cx = std::cos(pi * x[0]), sx = std::sin(pi * x[0]);
      double cy = std::cos(pi * x[1]), sy = std::sin(pi * x[1]);
      double cz = std::cos(pi * x[2]), sz = std::sin(pi * x[2]);
      double ct = std::cos(w * t), st = std::sin(w * t);
      const double ax = 1.0, ay = -1.0, az = 0.0;
      val[0] = ax * cx * sy * sz * ct;
      val[1] = ay * sx * cy * sz * ct;
      val[2] = az * sx * sy * cz * ct;
      val[3] = -pi * (az - ay) * sx * cy * cz / w * st;
      val[4] = -pi * (ax - az) * cx * sy * cz / w * st;
      val[5] = -pi * (ay - ax) * cx * cy * sz / w * st;
      return 0;
    }
  }
  return -1;
}

}  // namespace dgp
