/*
 * gp_dual.h -- forward-mode dual numbers for the adjoint step.
 *
 * Dual<K> carries a value and K directional derivatives. Instantiating the per-cell functions of gp_model.h with
 * R = Dual<K> evaluates a cell's outputs together with their derivatives along K seeded input directions: a block
 * of K columns of the cell's local Jacobian, exact to rounding (no finite differences). Branches follow the value
 * (comparisons look at .v), so the derivative is that of the branch the forward step takes; at the kinks of the
 * simplex projection and of |x| it is the one-sided derivative of that branch.
 */
#pragma once

#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define GD_HD __host__ __device__ __forceinline__
#else
#define GD_HD inline
#endif

namespace gpk {

/* the double forms stay visible next to the Dual overloads below */
using ::fabs;
using ::isfinite;
using ::sqrt;

template <int K>
struct Dual {
  double v;
  double d[K];
  GD_HD Dual() : v(0.0) {
    for (int i = 0; i < K; i++) d[i] = 0.0;
  }
  GD_HD Dual(double x) : v(x) {
    for (int i = 0; i < K; i++) d[i] = 0.0;
  }
  GD_HD explicit Dual(int x) : Dual((double)x) {}
  GD_HD Dual &operator+=(const Dual &o) {
    v += o.v;
    for (int i = 0; i < K; i++) d[i] += o.d[i];
    return *this;
  }
  GD_HD Dual &operator-=(const Dual &o) {
    v -= o.v;
    for (int i = 0; i < K; i++) d[i] -= o.d[i];
    return *this;
  }
  GD_HD Dual &operator*=(const Dual &o) {
    for (int i = 0; i < K; i++) d[i] = d[i] * o.v + v * o.d[i];
    v *= o.v;
    return *this;
  }
  GD_HD Dual &operator/=(const Dual &o) {
    const double q = v / o.v;
    for (int i = 0; i < K; i++) d[i] = (d[i] - q * o.d[i]) / o.v;
    v = q;
    return *this;
  }
};

template <int K> GD_HD Dual<K> operator+(Dual<K> a, const Dual<K> &b) { return a += b; }
template <int K> GD_HD Dual<K> operator-(Dual<K> a, const Dual<K> &b) { return a -= b; }
template <int K> GD_HD Dual<K> operator*(Dual<K> a, const Dual<K> &b) { return a *= b; }
template <int K> GD_HD Dual<K> operator/(Dual<K> a, const Dual<K> &b) { return a /= b; }
template <int K> GD_HD Dual<K> operator-(const Dual<K> &a) {
  Dual<K> r;
  r.v = -a.v;
  for (int i = 0; i < K; i++) r.d[i] = -a.d[i];
  return r;
}
template <int K> GD_HD bool operator==(const Dual<K> &a, const Dual<K> &b) { return a.v == b.v; }
template <int K> GD_HD bool operator!=(const Dual<K> &a, const Dual<K> &b) { return a.v != b.v; }
template <int K> GD_HD bool operator<(const Dual<K> &a, const Dual<K> &b) { return a.v < b.v; }
template <int K> GD_HD bool operator>(const Dual<K> &a, const Dual<K> &b) { return a.v > b.v; }
template <int K> GD_HD bool operator<=(const Dual<K> &a, const Dual<K> &b) { return a.v <= b.v; }
template <int K> GD_HD bool operator>=(const Dual<K> &a, const Dual<K> &b) { return a.v >= b.v; }
template <int K> GD_HD Dual<K> sqrt(const Dual<K> &a) {
  Dual<K> r;
  r.v = ::sqrt(a.v);
  const double g = 0.5 / r.v;
  for (int i = 0; i < K; i++) r.d[i] = a.d[i] * g;
  return r;
}
template <int K> GD_HD Dual<K> fabs(const Dual<K> &a) { return a.v < 0.0 ? -a : a; }
template <int K> GD_HD bool isfinite(const Dual<K> &a) { return ::isfinite(a.v); }

}  // namespace gpk
