// Conservation-law arithmetic shared by the host classes and the HIP kernels.
//
// Each law is a set of static inline functions templated on the arithmetic
// type T (double for the plain form, float for the reduced form).  No virtual
// dispatch reaches the device: kernels are instantiated per law.
//
//   ScalarLawD   ScalarLaw::Fi / riemannSolver / boundaryFlux   (ScalarLaw.cc:29-88)
//   EulerLawD    Euler2d::Fi (EulerLaw.cc:47-69), Euler2dRoe::riemannSolver
//                (EulerLaw.cc:913-1010), wall ghost (EulerLaw.cc:280-291), far-field
//                ghost (EulerLaw.cc:293-295), maximumEigenValue (EulerLaw.cc:1012-1034),
//                isPhysical (EulerLaw.cc:1398-1404), the exact Godunov flux of
//                Euler2d::riemannSolver (EulerLaw.cc:602-887); local Lax-Friedrichs
//                flux is new.
//   MaxwellLawD  3D Maxwell, eps = mu = 1, upwind/central flux, PEC ghost (new).
//   Maxwell2dLawD Maxwell2d, the TE fields (MaxwellLaw.cc:21-200).
//   Burgers2DLawD Burgers2D with its central flux (Burgers2D.cc:22-92).
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define DG_HD __host__ __device__ __forceinline__
#else
#define DG_HD inline
#endif

#include <math.h>

#include <type_traits>

namespace dgp {

enum { LAW_SCALAR = 0, LAW_EULER2D = 1, LAW_MAXWELL3D = 2, LAW_MAXWELL2D = 3, LAW_BURGERS2D = 4 };
enum { FLUX_UPWIND = 0, FLUX_LF = 1, FLUX_ROE = 2, FLUX_CENTRAL = 3, FLUX_ROE_EF = 4, FLUX_EXACT = 5 };
// boundary codes: far field (farfield[0..NF)), wall, PEC, second far field
// (farfield[4..8), Euler: a far-field state that differs by boundary, as the
// reference's far-field FieldEvaluator evaluated at the boundary point)
// BC_CHAR: the Euler characteristic far field (tag 30000), whose flux is the
// physical flux of a boundary state; the two-point flux takes that state on
// both sides (FP64 form)
// BC_HANG: a face of a mesh with hanging nodes (a coarse cell's face split on
// the other side, or a half of such a face): the stage kernels add nothing for
// it; the hanging-face pass integrates it over its halves
enum { BC_FARFIELD = 1, BC_WALL = 2, BC_PEC = 3, BC_FARFIELD2 = 4, BC_CHAR = 5, BC_HANG = 6 };

struct LawParams {
  double gamma;
  double vel[3];
  double farfield[8];
  int flux;
  int char_bc;  // 1: the mesh has characteristic far-field faces (BC_CHAR)
};

template <typename T>
DG_HD T dg_sqrt(T x) { return sqrt(x); }
template <>
DG_HD float dg_sqrt<float>(float x) { return sqrtf(x); }
template <typename T>
DG_HD T dg_abs(T x) { return x < T(0) ? -x : x; }
template <typename T>
DG_HD T dg_pow(T x, T y) { return pow(x, y); }
template <>
DG_HD float dg_pow<float>(float x, float y) { return powf(x, y); }
// reciprocal: exact division in FP64; the hardware reciprocal (1 ulp) in the
// FP32 form
template <typename T>
DG_HD T dg_rcp(T x) { return T(1) / x; }
template <>
DG_HD float dg_rcp<float>(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// ---------------------------------------------------------------------------
struct ScalarLawD {
  static constexpr int NF = 1;
  static constexpr int DIM = 2;
  static constexpr bool GHOST = true;  // boundary faces: ghost state + two-point flux
  DG_HD static constexpr bool both_sides(int) { return false; }
  template <typename T>
  DG_HD static void flux(const LawParams& L, const T* q, T F[][DIM]) {
    F[0][0] = T(L.vel[0]) * q[0];
    F[0][1] = T(L.vel[1]) * q[0];
  }
  template <typename T>
  DG_HD static void riemann(const LawParams& L, const T* n, const T* uL, const T* uR, T* out) {
    T vx = T(L.vel[0]), vy = T(L.vel[1]);
    T vn = n[0] * vx + n[1] * vy;
    T u = vn > T(0) ? uL[0] : uR[0];
    out[0] = (vx * u) * n[0] + (vy * u) * n[1];
  }
  // ghost state of a boundary face: the far-field value (ScalarLaw.cc:39-59 takes
  // the exterior value on inflow and the interior value on outflow, which is the
  // upwind flux against this ghost)
  template <typename T>
  DG_HD static void ghost(const LawParams& L, int code, const T* n, const T* q, T* g) {
    (void)code;
    (void)n;
    (void)q;
    g[0] = T(L.farfield[0]);
  }
  template <typename T>
  DG_HD static void boundary(const LawParams& L, int code, const T* n, const T* q, T* out) {
    T g[1];
    ghost(L, code, n, q, g);
    riemann(L, n, q, g, out);
  }
  template <typename T>
  DG_HD static bool physical(const LawParams&, const T*) { return true; }
  template <typename T>
  DG_HD static T maxev(const LawParams& L, const T*) {
    return T(fabs(L.vel[0]) + fabs(L.vel[1]) + fabs(L.vel[2]));
  }
};

// ScalarLaw with a velocity FieldEvaluator (ScalarLaw.cc:29-97, velocityField->eval
// at every point): the kernels read the velocity of each quadrature point from a
// table the host filled by evaluating the caller's field there (POINT_VEL), and
// pass it to flux_at / riemann_at; the far field and the ghost are ScalarLawD's
struct ScalarFieldLawD : ScalarLawD {
  static constexpr bool POINT_VEL = true;
  template <typename T>
  DG_HD static void flux_at(const T* v, const T* q, T F[][DIM]) {
    F[0][0] = v[0] * q[0];
    F[0][1] = v[1] * q[0];
  }
  // upwind by v.n at the point: ScalarLaw::riemannSolver; with the far-field
  // ghost on boundary faces it is ScalarLaw::boundaryFlux (inflow takes the far
  // field, outflow the interior)
  template <typename T>
  DG_HD static void riemann_at(const T* v, const T* n, const T* uL, const T* uR, T* out) {
    T vn = n[0] * v[0] + n[1] * v[1];
    T u = vn > T(0) ? uL[0] : uR[0];
    out[0] = (v[0] * u) * n[0] + (v[1] * u) * n[1];
  }
};

// true for laws whose flux needs the velocity of the point (ScalarFieldLawD)
template <class L, class = void>
struct point_vel : std::false_type {};
template <class L>
struct point_vel<L, std::void_t<decltype(L::POINT_VEL)>> : std::bool_constant<L::POINT_VEL> {};

// ---------------------------------------------------------------------------
struct EulerLawD {
  static constexpr int NF = 4;
  static constexpr int DIM = 2;
  static constexpr bool GHOST = true;  // boundary faces: ghost state + two-point flux
  template <typename T>
  DG_HD static void flux(const LawParams& L, const T* Q, T F[][DIM]) {
    T g = T(L.gamma);
    T invrho = dg_rcp(Q[0]);
    T Q1 = Q[1], Q2 = Q[2];
    T q12 = Q1 * Q2 * invrho, q11 = Q1 * Q1 * invrho, q22 = Q2 * Q2 * invrho;
    T p = (g - T(1)) * (Q[3] - T(0.5) * (q11 + q22));
    T qq = invrho * (Q[3] + p);
    F[0][0] = Q1;      F[0][1] = Q2;
    F[1][0] = q11 + p; F[1][1] = q12;
    F[2][0] = q12;     F[2][1] = q22 + p;
    F[3][0] = Q1 * qq; F[3][1] = Q2 * qq;
  }
  // F(u).n together with the normal velocity and the sound speed, sharing 1/rho
  template <typename T>
  DG_HD static void normal_flux(const LawParams& L, const T* n, const T* Q, T* Fn, T& vn, T& c) {
    const T g = T(L.gamma);
    const T ri = dg_rcp(Q[0]);
    const T mn = Q[1] * n[0] + Q[2] * n[1];
    const T p = (g - T(1)) * (Q[3] - T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2]) * ri);
    vn = mn * ri;
    c = dg_sqrt(g * p * ri);
    Fn[0] = mn;
    Fn[1] = Q[1] * vn + p * n[0];
    Fn[2] = Q[2] * vn + p * n[1];
    Fn[3] = (Q[3] + p) * vn;
  }
  template <typename T>
  DG_HD static T wave(const LawParams& L, const T* n, const T* u) {
    T g = T(L.gamma);
    T ri = T(1) / u[0];
    T p = (g - T(1)) * (u[3] - T(0.5) * (u[1] * u[1] + u[2] * u[2]) * ri);
    T c = dg_sqrt(g * p * ri);
    return c + dg_abs((u[1] * n[0] + u[2] * n[1]) * ri);
  }
  // local Lax-Friedrichs: (F(uL).n + F(uR).n)/2 - lam (uR - uL)/2 with
  // lam = max(|vn| + c) over both states (EulerLaw.cc:1022-1034)
  template <typename T>
  DG_HD static void lf(const LawParams& L, const T* n, const T* uL, const T* uR, T* out) {
    T FL[4], FR[4], vL, cL, vR, cR;
    normal_flux(L, n, uL, FL, vL, cL);
    normal_flux(L, n, uR, FR, vR, cR);
    const T l1 = cL + dg_abs(vL), l2 = cR + dg_abs(vR);
    const T lam = l1 > l2 ? l1 : l2;
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = T(0.5) * ((FL[i] + FR[i]) - lam * (uR[i] - uL[i]));
  }
  template <typename T>
  DG_HD static void roe(const LawParams& L, const T* n, const T* u0, const T* u1, T* flux_out) {
    T gamma = T(L.gamma);
    T rho_0 = u0[0], rho_1 = u1[0];
    T sqrt_rho_0 = dg_sqrt(rho_0), sqrt_rho_1 = dg_sqrt(rho_1);
    T inv_rho_0 = T(1) / rho_0, inv_rho_1 = T(1) / rho_1;
    T rho_A = sqrt_rho_0 * sqrt_rho_1;
    T inv_denom = T(1) / (sqrt_rho_0 + sqrt_rho_1);
    T u_0 = u0[1] * inv_rho_0, u_1 = u1[1] * inv_rho_1;
    T u_A = (u_0 * sqrt_rho_0 + u_1 * sqrt_rho_1) * inv_denom;
    T v_0 = u0[2] * inv_rho_0, v_1 = u1[2] * inv_rho_1;
    T v_A = (v_0 * sqrt_rho_0 + v_1 * sqrt_rho_1) * inv_denom;
    T p_0 = (gamma - T(1)) * (u0[3] - T(0.5) * (u0[1] * u0[1] + u0[2] * u0[2]) * inv_rho_0);
    T p_1 = (gamma - T(1)) * (u1[3] - T(0.5) * (u1[1] * u1[1] + u1[2] * u1[2]) * inv_rho_1);
    T H_0 = (u0[3] + p_0) * inv_rho_0, H_1 = (u1[3] + p_1) * inv_rho_1;
    T H_A = (H_0 * sqrt_rho_0 + H_1 * sqrt_rho_1) * inv_denom;
    T qq = H_A - T(0.5) * (u_A * u_A + v_A * v_A);
    T roe[4] = {T(0), T(0), T(0), T(0)};
    if (qq > T(0)) {
      T c_A = dg_sqrt((gamma - T(1)) * qq);
      T inv_c_A = T(1) / c_A;
      T vn = u_A * n[0] + v_A * n[1];
      T eg0 = dg_abs(vn), eg2 = dg_abs(vn + c_A), eg3 = dg_abs(vn - c_A);
      if (L.flux == FLUX_ROE_EF) {
        // Harten's entropy fix (Harten & Hyman 1983), delta = 0.2 c: a new option;
        // the reference's Roe flux has none, and its zero dissipation of the
        // linear waves on faces with vn ~ 0 lets a shock-tube run blow up
        const T dl = T(0.2) * c_A, inv2 = T(2.5) * inv_c_A;  // 1 / (2 delta)
        if (eg0 < dl) eg0 = (eg0 * eg0 + dl * dl) * inv2;
        if (eg2 < dl) eg2 = (eg2 * eg2 + dl * dl) * inv2;
        if (eg3 < dl) eg3 = (eg3 * eg3 + dl * dl) * inv2;
      }
      T k = T(0.5) * rho_A * inv_c_A;
      T drho = rho_1 - rho_0, dp = p_1 - p_0, du = u_1 - u_0, dv = v_1 - v_0;
      T ch0 = drho - dp * inv_c_A * inv_c_A;
      T ch1 = n[1] * du - n[0] * dv;
      T ch2 = n[0] * du + n[1] * dv + dp / rho_A * inv_c_A;
      T ch3 = -n[0] * du - n[1] * dv + dp / rho_A * inv_c_A;
      T a0 = eg0 * ch0, a1 = eg0 * ch1, a2 = eg2 * ch2, a3 = eg3 * ch3;
      // sum_i eval_i charv_i evec_i[j] with the eigenvectors of EulerLaw.cc:955-976
      roe[0] = a0 + a2 * k + a3 * k;
      roe[1] = a0 * u_A + a1 * (n[1] * rho_A) + a2 * (k * (u_A + c_A * n[0])) + a3 * (k * (u_A - c_A * n[0]));
      roe[2] = a0 * v_A + a1 * (-n[0] * rho_A) + a2 * (k * (v_A + c_A * n[1])) + a3 * (k * (v_A - c_A * n[1]));
      roe[3] = a0 * (T(0.5) * (u_A * u_A + v_A * v_A)) + a1 * (rho_A * (u_A * n[1] - v_A * n[0])) +
               a2 * (k * (H_A + c_A * vn)) + a3 * (k * (H_A - c_A * vn));
    }
    // F(u).n of both states from the velocities and pressures above:
    // (rho vn, rho u vn + p nx, rho v vn + p ny, (E + p) vn)
    const T vn0 = u_0 * n[0] + v_0 * n[1], vn1 = u_1 * n[0] + v_1 * n[1];
    const T fn[4] = {u0[0] * vn0 + u1[0] * vn1, u0[1] * vn0 + u1[1] * vn1 + (p_0 + p_1) * n[0],
                     u0[2] * vn0 + u1[2] * vn1 + (p_0 + p_1) * n[1], (u0[3] + p_0) * vn0 + (u1[3] + p_1) * vn1};
#pragma unroll
    for (int i = 0; i < 4; i++) flux_out[i] = T(0.5) * (fn[i] - roe[i]);
  }
  // Exact Riemann problem in the face frame (u normal, v tangential velocity),
  // sampled at x/t = 0, in the steps of Euler::riemann (EulerLaw.cc:602-851):
  // both states supersonic one way -> that state; else Newton for p* from the
  // two-shock guess (at most 12 iterations while the relative change is
  // >= 1e-14, p* floored at 1e-14), u* from the final p*, the side by the sign
  // of u*, and on that side the star state for a shock (no shock-speed test),
  // the star state when the rarefaction tail and u* have opposite signs, else
  // the fan state at x/t = 0 (no fan-head test).
  template <typename T>
  DG_HD static void exact_state(T g, T rl, T ul, T vl, T pl, T rr, T ur, T vr, T pr, T& rho, T& u, T& v, T& p) {
    const T TOL = T(1e-14), gm1 = g - T(1), gp1 = g + T(1);
    const T ril = T(1) / rl, rir = T(1) / rr;
    const T cl = dg_sqrt(dg_abs(g * pl * ril)), cr = dg_sqrt(dg_abs(g * pr * rir));
    if (ul >= cl && ur >= cr) {
      rho = rl, u = ul, v = vl, p = pl;
      return;
    }
    if (ur <= -cr && ul <= -cl) {
      rho = rr, u = ur, v = vr, p = pr;
      return;
    }
    const T ig = T(1) / g, C3 = T(0.5) * gm1 * ig, C4 = T(2) / gm1;
    const T du = ur - ul;
    T p0 = T(0.5) * (pl + pr) - T(0.125) * du * (rl + rr) * (cl + cr);
    p0 = p0 < TOL ? TOL : p0;
    const T gl = dg_sqrt(T(2) / (rl * (gm1 * pl + gp1 * p0)));
    const T gr = dg_sqrt(T(2) / (rr * (gm1 * pr + gp1 * p0)));
    T ps = (gl * pl + gr * pr - du) / (gl + gr);
    ps = ps < TOL ? TOL : ps;
    T prev = ps, rel = TOL;
    // f_K(p) and its derivative: shock branch above p_K, rarefaction below
    auto branch = [&](T pk, T rik, T ck, T& f, T& df) {
      if (ps > pk) {
        const T a = T(1) / (gp1 * ps + gm1 * pk);
        const T b = dg_sqrt(T(2) * a * rik);
        f = (ps - pk) * b;
        df = b * (T(1) - T(0.5) * gp1 * (ps - pk) * a);
      } else {
        const T b = dg_pow(ps / pk, C3);
        f = C4 * ck * (b - T(1));
        df = b * ck * ig / ps;
      }
    };
    for (int it = 0; it < 12 && rel >= TOL; it++) {
      T fl, dfl, fr, dfr;
      branch(pl, ril, cl, fl, dfl);
      branch(pr, rir, cr, fr, dfr);
      ps -= (fl + fr + du) / (dfl + dfr);
      ps = ps < TOL ? TOL : ps;
      rel = dg_abs(ps - prev) / (T(0.5) * (prev + ps));
      prev = ps;
    }
    const T fl = ps > pl ? (ps - pl) * dg_sqrt(T(2) * ril / (gp1 * ps + gm1 * pl)) : C4 * cl * (dg_pow(ps / pl, C3) - T(1));
    const T fr = ps > pr ? (ps - pr) * dg_sqrt(T(2) * rir / (gp1 * ps + gm1 * pr)) : C4 * cr * (dg_pow(ps / pr, C3) - T(1));
    const T us = (ul + ur + fr - fl) * T(0.5);
    const bool left = us >= T(0);
    const T pK = left ? pl : pr, uK = left ? ul : ur, rK = left ? rl : rr, cK = left ? cl : cr;
    const T sg = left ? T(1) : T(-1);
    v = left ? vl : vr;
    if (ps >= pK) {
      const T C = gm1 / gp1, q = ps / pK;
      rho = rK * (q + C) / (C * q + T(1));
      u = us;
      p = ps;
      return;
    }
    const T rstar = rK * dg_pow(ps / pK, ig);
    const T cstar = dg_sqrt(g * ps / rstar);
    if ((us - sg * cstar) * us <= T(0)) {
      rho = rstar, u = us, p = ps;
      return;
    }
    const T C = T(2) / gp1, q = C + sg * gm1 * uK / (cK * gp1);
    rho = rK * dg_pow(q, C4);
    u = C * (gm1 * T(0.5) * uK + sg * cK);
    p = pK * dg_pow(q, C4 * g);
  }
  // exact Godunov flux (Euler2d::riemannSolver, EulerLaw.cc:853-887, the
  // reference's default Euler flux): the exact state turned back to (x, y), its
  // physical flux dotted with n
  template <typename T>
  DG_HD static void godunov(const LawParams& L, const T* n, const T* u0, const T* u1, T* out) {
    const T g = T(L.gamma), gm1 = g - T(1);
    const T vn1 = n[0] * (u0[1] / u0[0]) + n[1] * (u0[2] / u0[0]);
    const T vt1 = n[1] * (u0[1] / u0[0]) - n[0] * (u0[2] / u0[0]);
    const T vn2 = n[0] * (u1[1] / u1[0]) + n[1] * (u1[2] / u1[0]);
    const T vt2 = n[1] * (u1[1] / u1[0]) - n[0] * (u1[2] / u1[0]);
    const T p1 = gm1 * u0[3] - T(0.5) * gm1 * (u0[1] * u0[1] + u0[2] * u0[2]) / u0[0];
    const T p2 = gm1 * u1[3] - T(0.5) * gm1 * (u1[1] * u1[1] + u1[2] * u1[2]) / u1[0];
    T r, vn, vt, p;
    exact_state(g, u0[0], vn1, vt1, p1, u1[0], vn2, vt2, p2, r, vn, vt, p);
    const T q[4] = {r, r * (n[0] * vn + n[1] * vt), r * (n[1] * vn - n[0] * vt),
                    (T(1) / gm1) * (p + T(0.5) * gm1 * r * (vn * vn + vt * vt))};
    T F[4][DIM];
    flux(L, q, F);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = F[i][0] * n[0] + F[i][1] * n[1];
  }
  template <typename T>
  DG_HD static void riemann(const LawParams& L, const T* n, const T* uL, const T* uR, T* out) {
    if (L.flux == FLUX_ROE || L.flux == FLUX_ROE_EF)
      roe(L, n, uL, uR, out);
    else
      lf(L, n, uL, uR, out);
  }
  // ghost state of a boundary face (EulerLaw.cc:280-295): the wall mirrors the
  // normal momentum, every other tag takes the far-field state
  template <typename T>
  DG_HD static void ghost(const LawParams& L, int code, const T* n, const T* Q, T* g) {
    if (code == BC_WALL) {
      T vn = -(Q[1] * n[0] + Q[2] * n[1]);
      T vt = Q[1] * n[1] - Q[2] * n[0];
      g[0] = Q[0];
      g[1] = vn * n[0] + vt * n[1];
      g[2] = vn * n[1] - vt * n[0];
      g[3] = Q[3];
    } else {
      const int o = code == BC_FARFIELD2 ? 4 : 0;
      g[0] = T(L.farfield[o + 0]);
      g[1] = T(L.farfield[o + 1]);
      g[2] = T(L.farfield[o + 2]);
      g[3] = T(L.farfield[o + 3]);
    }
  }
  template <typename T>
  DG_HD static void boundary(const LawParams& L, int code, const T* n, const T* Q, T* out) {
    T g[4];
    ghost(L, code, n, Q, g);
    riemann(L, n, Q, g, out);
  }
  DG_HD static constexpr bool both_sides(int) { return false; }
  // the boundary state of the characteristic far field (Euler2d::boundaryFlux,
  // tag 30000, EulerLaw.cc:187-278): supersonic inflow the far field, supersonic
  // outflow the interior state, subsonic the state of the Riemann invariants
  // J+ and J- taken from the far field and the interior
  template <typename T>
  DG_HD static void char_state(const LawParams& L, const T* n, const T* Q, T* g) {
    const T gamma = T(L.gamma), gm1 = gamma - T(1), ogm1 = T(1) / gm1;
    const T qi[4] = {T(L.farfield[0]), T(L.farfield[1]), T(L.farfield[2]), T(L.farfield[3])};
    const T inv_rho = T(1) / Q[0], inv_rho_inf = T(1) / qi[0];
    const T p = gm1 * (Q[3] - T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2]) * inv_rho);
    const T pinf = gm1 * (qi[3] - T(0.5) * (qi[1] * qi[1] + qi[2] * qi[2]) * inv_rho_inf);
    const T vn = (n[0] * Q[1] + n[1] * Q[2]) * inv_rho;
    const T vn_inf = (n[0] * qi[1] + n[1] * qi[2]) * inv_rho_inf;
    const T c = dg_sqrt(gamma * p * inv_rho);
    const bool inflow = vn < T(0);
    if (dg_abs(vn) >= c) {
      for (int i = 0; i < 4; i++) g[i] = inflow ? qi[i] : Q[i];
      return;
    }
    const T c_inf = dg_sqrt(gamma * pinf * inv_rho_inf);
    T vn_b, c_b, vt, r_b;
    if (inflow) {
      const T jp = -vn_inf + T(2) * c_inf * ogm1, jm = -vn - c * T(2) * ogm1;
      vn_b = T(-0.5) * (jp + jm);
      c_b = T(0.25) * (jp - jm) * gm1;
      vt = (n[1] * qi[1] - n[0] * qi[2]) * inv_rho_inf;
      r_b = qi[0] * dg_pow(c_b / c_inf, T(2) * ogm1);
    } else {
      const T jm = vn_inf - T(2) * c_inf * ogm1, jp = vn + T(2) * c * ogm1;
      vn_b = T(0.5) * (jp + jm);
      c_b = T(0.25) * (jp - jm) * gm1;
      vt = (n[1] * Q[1] - n[0] * Q[2]) * inv_rho;
      r_b = Q[0] * dg_pow(c_b / c, T(2) * ogm1);
    }
    g[0] = r_b;
    g[1] = r_b * (n[0] * vn_b + n[1] * vt);
    g[2] = r_b * (n[1] * vn_b - n[0] * vt);
    const T p_b = r_b * c_b * c_b / gamma;
    g[3] = p_b * ogm1 + T(0.5) * (g[1] * g[1] + g[2] * g[2]) / r_b;
  }
  template <typename T>
  DG_HD static bool physical(const LawParams& L, const T* Q) {
    if (Q[0] < T(0)) return false;
    if ((T(L.gamma) - T(1)) * (Q[3] - T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2]) / Q[0]) < T(0)) return false;
    return true;
  }
  template <typename T>
  DG_HD static T maxev(const LawParams& L, const T* Q) {
    T rho_inv = T(1) / Q[0];
    T v2 = (Q[1] * Q[1] + Q[2] * Q[2]) * rho_inv;
    T p = (T(L.gamma) - T(1)) * (Q[3] - T(0.5) * v2);
    v2 *= rho_inv;
    T c = dg_sqrt(T(L.gamma) * p * rho_inv);
    return c + dg_sqrt(v2);
  }
};

// Euler2d with its own numerical flux, the exact Godunov flux: a law type of
// its own, so the kernels of the other Euler fluxes keep their register budget
// (the Newton solver would otherwise set it for all of them)
struct EulerExactLawD : EulerLawD {
  template <typename T>
  DG_HD static void riemann(const LawParams& L, const T* n, const T* uL, const T* uR, T* out) {
    godunov(L, n, uL, uR, out);
  }
  template <typename T>
  DG_HD static void boundary(const LawParams& L, int code, const T* n, const T* Q, T* out) {
    T g[4];
    ghost(L, code, n, Q, g);
    godunov(L, n, Q, g, out);
  }
};

// The Euler laws on meshes with characteristic far-field faces (tag 30000,
// BC_CHAR; FP64 form): their ghost on those faces is the boundary state, taken
// on both sides of the two-point flux, so the flux is that state's physical
// flux (consistency of every Euler flux).  Types of their own: the other Euler
// kernels do not carry the branch.
template <class BASE>
struct EulerCharLaw : BASE {
  DG_HD static constexpr bool both_sides(int code) { return code == BC_CHAR; }
  template <typename T>
  DG_HD static void ghost(const LawParams& L, int code, const T* n, const T* Q, T* g) {
    if (code == BC_CHAR)
      BASE::char_state(L, n, Q, g);
    else
      BASE::ghost(L, code, n, Q, g);
  }
  template <typename T>
  DG_HD static void boundary(const LawParams& L, int code, const T* n, const T* Q, T* out) {
    T g[4];
    ghost(L, code, n, Q, g);
    BASE::riemann(L, n, code == BC_CHAR ? g : Q, g, out);
  }
};
using EulerCharLawD = EulerCharLaw<EulerLawD>;
using EulerExactCharLawD = EulerCharLaw<EulerExactLawD>;

// Maxwell2d (MaxwellLaw.cc:21-37, 115-200): the TE fields (Hx, Hy, Ez), eps = mu
// = 1.  F_Hx = (0, Ez), F_Hy = (-Ez, 0), F_Ez = (-Hy, Hx).  The numerical flux is
// the physical flux of the reference's interface state, in its condensed form
// (MaxwellLaw.cc:170-175).  Walls (40000, and 20000 on straight faces) mirror H
// about the normal (MaxwellLaw.cc:37-56, 115-126); other boundary faces take
// the far-field state; the wave speed is 1.  The reference's RHS holds a pulse
// source of one scenario (MaxwellLaw.cc:362-367); this law has none.
struct Maxwell2dLawD {
  static constexpr int NF = 3;
  static constexpr int DIM = 2;
  static constexpr bool GHOST = true;
  DG_HD static constexpr bool both_sides(int) { return false; }
  template <typename T>
  DG_HD static void flux(const LawParams&, const T* Q, T F[][DIM]) {
    F[0][0] = T(0);  F[0][1] = Q[2];
    F[1][0] = -Q[2]; F[1][1] = T(0);
    F[2][0] = -Q[1]; F[2][1] = Q[0];
  }
  template <typename T>
  DG_HD static void riemann(const LawParams&, const T* n, const T* u0, const T* u1, T* out) {
    const T nx = n[0], ny = n[1], h = T(2);
    const T s0 = (u0[0] + u1[0]) * (nx * ny / h + ny * ny / h) + (u0[1] + u1[1]) * (nx * nx / h - nx * ny / h) +
                 (u0[2] - u1[2]) * ny / h;
    const T s1 = (u0[0] + u1[0]) * (-nx * ny / h + ny * ny / h) + (u0[1] + u1[1]) * (nx * nx / h + nx * ny / h) +
                 (u0[2] - u1[2]) * (-nx) / h;
    const T s2 = ny * (u0[0] - u1[0]) / h + nx * (u1[1] - u0[1]) / h + (u0[2] + u1[2]) / h;
    out[0] = s2 * ny;
    out[1] = -s2 * nx;
    out[2] = -s1 * nx + s0 * ny;
  }
  template <typename T>
  DG_HD static void ghost(const LawParams& L, int code, const T* n, const T* Q, T* g) {
    if (code == BC_WALL) {
      const T hn = Q[0] * n[0] + Q[1] * n[1];
      g[0] = -Q[0] + T(2) * n[0] * hn;
      g[1] = -Q[1] + T(2) * n[1] * hn;
      g[2] = Q[2];
    } else {
      for (int i = 0; i < 3; i++) g[i] = T(L.farfield[i]);
    }
  }
  template <typename T>
  DG_HD static void boundary(const LawParams& L, int code, const T* n, const T* Q, T* out) {
    T g[3];
    ghost(L, code, n, Q, g);
    riemann(L, n, Q, g, out);
  }
  template <typename T>
  DG_HD static bool physical(const LawParams&, const T*) { return true; }
  template <typename T>
  DG_HD static T maxev(const LawParams&, const T*) { return T(1); }
};

// Burgers2D (Burgers2D.cc:22-92): F = (u^2/2, u^2/2).  The numerical flux is the
// reference's central flux (its dissipation coefficient maxIV is 0, :74-77).  A
// boundary face takes the interior flux on outflow and the far-field flux on
// inflow (boundaryFlux, :31-49); no ghost state gives that through a central
// flux, so the law supplies its boundary flux (GHOST = false).  The wave speed
// is 1 (maximumEigenValue, :88-92).
struct Burgers2DLawD {
  static constexpr int NF = 1;
  static constexpr int DIM = 2;
  static constexpr bool GHOST = false;
  DG_HD static constexpr bool both_sides(int) { return false; }
  template <typename T>
  DG_HD static void flux(const LawParams&, const T* Q, T F[][DIM]) {
    F[0][0] = Q[0] * Q[0] * T(0.5);
    F[0][1] = Q[0] * Q[0] * T(0.5);
  }
  template <typename T>
  DG_HD static void riemann(const LawParams&, const T* n, const T* u0, const T* u1, T* out) {
    out[0] = T(0.5) * ((u0[0] * u0[0] + u1[0] * u1[0]) * (n[0] + n[1]) * T(0.5));
  }
  // the state seen across a boundary face by the limiter: the far field
  template <typename T>
  DG_HD static void ghost(const LawParams& L, int, const T*, const T*, T* g) {
    g[0] = T(L.farfield[0]);
  }
  template <typename T>
  DG_HD static void boundary(const LawParams& L, int, const T* n, const T* Q, T* out) {
    const T vn = Q[0] * Q[0] * T(0.5) * (n[0] + n[1]);
    const T ff = T(L.farfield[0]);
    out[0] = vn < T(0) ? ff * ff * T(0.5) * (n[0] + n[1]) : vn;
  }
  template <typename T>
  DG_HD static bool physical(const LawParams&, const T*) { return true; }
  template <typename T>
  DG_HD static T maxev(const LawParams&, const T*) { return T(1); }
};

// Numerical flux on a face.  Interior faces (link >= 0): the law's two-point
// flux of the two traces.  Boundary faces (link < 0, code -link): the two-point
// flux against the law's ghost state, written over uR, or the law's own
// boundary flux when no ghost state expresses it.
template <class LAW, typename T>
DG_HD void face_flux(const LawParams& L, int link, const T* n, const T* uL, T* uR, T* out) {
  if constexpr (LAW::GHOST) {
    if (link < 0) LAW::ghost(L, -link, n, uL, uR);
    if constexpr (LAW::both_sides(BC_CHAR)) {
      // a boundary state for both sides (the Euler characteristic far field)
      const bool both = link < 0 && LAW::both_sides(-link);
      T ul[LAW::NF];
#pragma unroll
      for (int f = 0; f < LAW::NF; f++) ul[f] = both ? uR[f] : uL[f];
      LAW::riemann(L, n, ul, uR, out);
    } else {
      LAW::riemann(L, n, uL, uR, out);
    }
  } else {
    if (link < 0)
      LAW::boundary(L, -link, n, uL, out);
    else
      LAW::riemann(L, n, uL, uR, out);
  }
}

// ---------------------------------------------------------------------------
// Maxwell in 3D with eps = mu = 1: dE/dt = curl H, dH/dt = -curl E, written
// as dq/dt + div F(q) = 0 with (F.n)_E = -n x H and (F.n)_H = n x E.
struct MaxwellLawD {
  static constexpr int NF = 6;
  static constexpr int DIM = 3;
  static constexpr bool GHOST = true;  // boundary faces: ghost state + two-point flux
  DG_HD static constexpr bool both_sides(int) { return false; }
  template <typename T>
  DG_HD static void flux(const LawParams&, const T* Q, T F[][DIM]) {
    const T *E = Q, *H = Q + 3;
    F[0][0] = T(0);  F[0][1] = -H[2]; F[0][2] = H[1];
    F[1][0] = H[2];  F[1][1] = T(0);  F[1][2] = -H[0];
    F[2][0] = -H[1]; F[2][1] = H[0];  F[2][2] = T(0);
    F[3][0] = T(0);  F[3][1] = E[2];  F[3][2] = -E[1];
    F[4][0] = -E[2]; F[4][1] = T(0);  F[4][2] = E[0];
    F[5][0] = E[1];  F[5][1] = -E[0]; F[5][2] = T(0);
  }
  template <typename T>
  DG_HD static void riemann(const LawParams& L, const T* n, const T* uL, const T* uR, T* out) {
    T Ea[3], Ha[3], dE[3], dH[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      Ea[i] = T(0.5) * (uL[i] + uR[i]);
      Ha[i] = T(0.5) * (uL[3 + i] + uR[3 + i]);
      dE[i] = uR[i] - uL[i];
      dH[i] = uR[3 + i] - uL[3 + i];
    }
    out[0] = -(n[1] * Ha[2] - n[2] * Ha[1]);
    out[1] = -(n[2] * Ha[0] - n[0] * Ha[2]);
    out[2] = -(n[0] * Ha[1] - n[1] * Ha[0]);
    out[3] = n[1] * Ea[2] - n[2] * Ea[1];
    out[4] = n[2] * Ea[0] - n[0] * Ea[2];
    out[5] = n[0] * Ea[1] - n[1] * Ea[0];
    if (L.flux == FLUX_UPWIND) {
      T ndE = n[0] * dE[0] + n[1] * dE[1] + n[2] * dE[2];
      T ndH = n[0] * dH[0] + n[1] * dH[1] + n[2] * dH[2];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        out[i] += T(0.5) * (n[i] * ndE - dE[i]);
        out[3 + i] += T(0.5) * (n[i] * ndH - dH[i]);
      }
    }
  }
  // perfect electric conductor: E -> -E, H -> H in the ghost
  template <typename T>
  DG_HD static void ghost(const LawParams& L, int code, const T* n, const T* Q, T* g) {
    (void)L;
    (void)code;
    (void)n;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      g[i] = -Q[i];
      g[3 + i] = Q[3 + i];
    }
  }
  template <typename T>
  DG_HD static void boundary(const LawParams& L, int code, const T* n, const T* Q, T* out) {
    T g[6];
    ghost(L, code, n, Q, g);
    riemann(L, n, Q, g, out);
  }
  template <typename T>
  DG_HD static bool physical(const LawParams&, const T*) { return true; }
  template <typename T>
  DG_HD static T maxev(const LawParams&, const T*) { return T(1); }
};

}  // namespace dgp
