// Forward-mode dual numbers with N tangent directions, used to differentiate
// the pointwise flux functions (physical flux, numerical flux, boundary flux)
// inside the tangent-linear and adjoint kernels.  The flux functions are
// templates on their arithmetic type, so the same source that computes the
// residual computes its derivatives; branches (max of wave speeds, upwinding,
// the Roe positivity test) follow the value part.
#pragma once

#include "dg_laws.h"

namespace dgp {

template <typename T, int N>
struct Dual {
  T v;
  T d[N];
  DG_HD Dual() : v(T(0)) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = T(0);
  }
  DG_HD Dual(T x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = T(0);
  }
  DG_HD explicit Dual(int x) : Dual(T(x)) {}
  DG_HD Dual& operator+=(const Dual& o) {
    v += o.v;
#pragma unroll
    for (int i = 0; i < N; i++) d[i] += o.d[i];
    return *this;
  }
  DG_HD Dual& operator-=(const Dual& o) {
    v -= o.v;
#pragma unroll
    for (int i = 0; i < N; i++) d[i] -= o.d[i];
    return *this;
  }
  DG_HD Dual& operator*=(const Dual& o) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = d[i] * o.v + v * o.d[i];
    v *= o.v;
    return *this;
  }
};

template <typename T, int N>
DG_HD Dual<T, N> operator+(Dual<T, N> a, const Dual<T, N>& b) {
  return a += b;
}
template <typename T, int N>
DG_HD Dual<T, N> operator-(Dual<T, N> a, const Dual<T, N>& b) {
  return a -= b;
}
template <typename T, int N>
DG_HD Dual<T, N> operator-(const Dual<T, N>& a) {
  Dual<T, N> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <typename T, int N>
DG_HD Dual<T, N> operator*(Dual<T, N> a, const Dual<T, N>& b) {
  return a *= b;
}
template <typename T, int N>
DG_HD Dual<T, N> operator/(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r;
  const T inv = T(1) / b.v;
  r.v = a.v * inv;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
  return r;
}
// mixed operations with a plain value (tables, constants)
template <typename T, int N>
DG_HD Dual<T, N> operator*(const T& a, Dual<T, N> b) {
  b.v *= a;
#pragma unroll
  for (int i = 0; i < N; i++) b.d[i] *= a;
  return b;
}
template <typename T, int N>
DG_HD Dual<T, N> operator*(Dual<T, N> b, const T& a) {
  return a * b;
}

template <typename T, int N>
DG_HD bool operator<(const Dual<T, N>& a, const Dual<T, N>& b) {
  return a.v < b.v;
}
template <typename T, int N>
DG_HD bool operator>(const Dual<T, N>& a, const Dual<T, N>& b) {
  return a.v > b.v;
}

template <typename T, int N>
DG_HD bool operator<=(const Dual<T, N>& a, const Dual<T, N>& b) {
  return a.v <= b.v;
}
template <typename T, int N>
DG_HD bool operator>=(const Dual<T, N>& a, const Dual<T, N>& b) {
  return a.v >= b.v;
}

// a^b: d(a^b) = b a^(b-1) da + a^b ln(a) db (the second term only where b
// carries a tangent: the exponents of the Riemann solver are constants)
template <typename T, int N>
DG_HD Dual<T, N> dg_pow(Dual<T, N> a, Dual<T, N> b) {
  Dual<T, N> r;
  r.v = dg_pow(a.v, b.v);
  const T da = b.v * dg_pow(a.v, b.v - T(1));
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = da * a.d[i] + (b.d[i] != T(0) ? r.v * log(a.v) * b.d[i] : T(0));
  return r;
}

template <typename T, int N>
DG_HD Dual<T, N> dg_sqrt(Dual<T, N> a) {
  Dual<T, N> r;
  r.v = dg_sqrt(a.v);
  const T h = r.v > T(0) ? T(0.5) / r.v : T(0);
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * h;
  return r;
}
template <typename T, int N>
DG_HD Dual<T, N> dg_abs(Dual<T, N> a) {
  return a.v < T(0) ? -a : a;
}

}  // namespace dgp
