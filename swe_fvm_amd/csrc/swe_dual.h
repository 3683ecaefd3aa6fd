// Forward-mode dual numbers for the adjoint: Dual<N> carries a value and
// its derivatives along N input directions.  The flux functions are
// templates on their scalar type; instantiated on Dual<N> they return the
// exact local Jacobian of the same operation sequence (branch decisions
// and min / max selections are taken on the values, as in the forward
// step), which the reverse pass contracts with the flux adjoint.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace swe {

// the Dual overloads below must not hide the double ones
using ::fabs;
using ::sqrt;

template <int N>
struct Dual {
  double v;
  double d[N];
};

template <int N>
__host__ __device__ __forceinline__ Dual<N> dconst(double v) {
  Dual<N> r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = 0.;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> dvar(double v, int k) {
  Dual<N> r = dconst<N>(v);
  r.d[k] = 1.;
  return r;
}

// a zero of the scalar type (positive zero, no derivatives)
template <class T>
struct Zero;
template <>
struct Zero<double> {
  __host__ __device__ static double get() { return 0.; }
};
template <int N>
struct Zero<Dual<N>> {
  __host__ __device__ static Dual<N> get() { return dconst<N>(0.); }
};

__host__ __device__ __forceinline__ double val(double x) { return x; }
template <int N>
__host__ __device__ __forceinline__ double val(const Dual<N>& x) { return x.v; }

#define SWE_DUAL_BINOP(OP, VEXPR, DEXPR)                                                   \
  template <int N>                                                                         \
  __host__ __device__ __forceinline__ Dual<N> operator OP(const Dual<N>& a, const Dual<N>& b) { \
    Dual<N> r;                                                                             \
    r.v = VEXPR;                                                                           \
    _Pragma("unroll") for (int k = 0; k < N; ++k) r.d[k] = DEXPR;                          \
    return r;                                                                              \
  }
SWE_DUAL_BINOP(+, a.v + b.v, a.d[k] + b.d[k])
SWE_DUAL_BINOP(-, a.v - b.v, a.d[k] - b.d[k])
SWE_DUAL_BINOP(*, a.v * b.v, a.d[k] * b.v + a.v * b.d[k])
#undef SWE_DUAL_BINOP

template <int N>
__host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  // the value is the exact quotient (branch decisions downstream compare
  // it); the derivatives use one reciprocal
  Dual<N> r;
  r.v = a.v / b.v;
  const double ib = 1. / b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
// mixed with plain doubles (constants of the step: geometry, tolerances)
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) {
  Dual<N> r = b;
  r.v = a + b.v;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v / b;
  const double ib = 1. / b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * ib;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) {
  return dconst<N>(a) / b;
}
#define SWE_DUAL_CMP(OP)                                                                     \
  template <int N>                                                                           \
  __host__ __device__ __forceinline__ bool operator OP(const Dual<N>& a, const Dual<N>& b) { \
    return a.v OP b.v;                                                                       \
  }                                                                                          \
  template <int N>                                                                           \
  __host__ __device__ __forceinline__ bool operator OP(const Dual<N>& a, double b) {         \
    return a.v OP b;                                                                         \
  }                                                                                          \
  template <int N>                                                                           \
  __host__ __device__ __forceinline__ bool operator OP(double a, const Dual<N>& b) {         \
    return a OP b.v;                                                                         \
  }
SWE_DUAL_CMP(<)
SWE_DUAL_CMP(<=)
SWE_DUAL_CMP(>)
SWE_DUAL_CMP(>=)
#undef SWE_DUAL_CMP

template <int N>
__host__ __device__ __forceinline__ Dual<N> sqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = ::sqrt(a.v);
  const double g = r.v > 0. ? 0.5 / r.v : 0.;  // the derivative at 0 is taken as 0
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * g;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> fabs(const Dual<N>& a) {
  return a.v < 0. ? -a : a;
}
// std::min / std::max semantics on the values
template <int N>
__host__ __device__ __forceinline__ Dual<N> dmin(const Dual<N>& a, const Dual<N>& b) { return (b < a) ? b : a; }
template <int N>
__host__ __device__ __forceinline__ Dual<N> dmax(const Dual<N>& a, const Dual<N>& b) { return (a < b) ? b : a; }
template <int N>
__host__ __device__ __forceinline__ Dual<N> dmin(double a, const Dual<N>& b) { return dmin(dconst<N>(a), b); }
template <int N>
__host__ __device__ __forceinline__ Dual<N> dmax(double a, const Dual<N>& b) { return dmax(dconst<N>(a), b); }
template <int N>
__host__ __device__ __forceinline__ bool is_wet(const Dual<N>& h) { return h.v > 1e-12; }

}  // namespace swe
