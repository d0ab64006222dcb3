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

// Exact solution of Sod's Riemann problem at x/t = s (Toro, "Riemann Solvers
// and Numerical Methods for Fluid Dynamics", ch. 4): the star pressure from
// Newton on the pressure function, then the wave structure of Sod's data --
// a rarefaction to the left, contact, shock to the right.  out: (rho, u, p).
static void sod_exact(double g, double s, double* out) {
  const double rl = 1.0, pl = 1.0, rr = 0.125, pr = 0.1;  // Examples.cc:62-83, u = 0 on both sides
  const double cl = std::sqrt(g * pl / rl), cr = std::sqrt(g * pr / rr);
  const double ar = 2.0 / ((g + 1.0) * rr), br = (g - 1.0) / (g + 1.0) * pr;
  const double ex = (g - 1.0) / (2.0 * g);
  // f(p) = f_L(p) + f_R(p); Sod's star pressure lies between pr and pl, so the
  // left branch is the rarefaction and the right one the shock
  double p = 0.5 * (pl + pr);
  for (int it = 0; it < 100; it++) {
    const double fl = 2.0 * cl / (g - 1.0) * (std::pow(p / pl, ex) - 1.0);
    const double dfl = std::pow(p / pl, -(g + 1.0) / (2.0 * g)) / (rl * cl);
    const double qr = std::sqrt(ar / (p + br));
    const double fr = (p - pr) * qr, dfr = qr * (1.0 - 0.5 * (p - pr) / (p + br));
    const double pn = p - (fl + fr) / (dfl + dfr);
    const double ch = std::fabs(pn - p) / (0.5 * (pn + p));
    p = pn;
    if (ch < 1e-15) break;
  }
  const double us = (p - pr) * std::sqrt(ar / (p + br));  // u* = u_R + f_R(p*)
  if (s >= us) {
    const double sh = cr * std::sqrt((g + 1.0) / (2.0 * g) * p / pr + ex);  // shock speed (u_R = 0)
    if (s >= sh) {
      out[0] = rr, out[1] = 0.0, out[2] = pr;
    } else {
      const double k = (g - 1.0) / (g + 1.0), q = p / pr;
      out[0] = rr * (q + k) / (k * q + 1.0), out[1] = us, out[2] = p;
    }
    return;
  }
  const double cs = cl * std::pow(p / pl, ex);
  if (s <= -cl) {
    out[0] = rl, out[1] = 0.0, out[2] = pl;
  } else if (s >= us - cs) {
    out[0] = rl * std::pow(p / pl, 1.0 / g), out[1] = us, out[2] = p;
  } else {  // inside the left fan: u - c = s, u + 2c/(g-1) = 2 c_L/(g-1)
    const double c = 2.0 / (g + 1.0) * (cl - 0.5 * (g - 1.0) * s);
    out[0] = rl * std::pow(c / cl, 2.0 / (g - 1.0));
    out[1] = 2.0 / (g + 1.0) * (cl + s);
    out[2] = pl * std::pow(c / cl, 2.0 * g / (g - 1.0));
  }
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
      // ShockTube2 (Examples.cc:62-83), interface x = par[0]: the initial step at
      // t = 0, the exact solution of the Riemann problem at t > 0
      double w[3];
      if (t > 0.0) {
        sod_exact(law.gamma, (x[0] - par[0]) / t, w);
      } else {
        const bool left = x[0] < par[0];
        w[0] = left ? 1.0 : 0.125, w[1] = 0.0, w[2] = left ? 1.0 : 0.1;
      }
      val[0] = w[0];
      val[1] = w[0] * w[1];
      val[2] = 0.0;
      val[3] = w[2] / (law.gamma - 1.0) + 0.5 * w[0] * w[1] * w[1];
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
