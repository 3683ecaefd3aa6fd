// The partially wet reconstructions on a generic scalar type, for the
// derivative builds (tangent-linear kernels on Dual<1>, the adjoint's local
// Jacobians on Dual<N>): ReconstructPartWetCell1 (src/MUSCLObject.cpp:86-112)
// and ReconstructPartWetCell2 (src/MUSCLObject.cpp:114-191, with the
// write-row decision D4 of the step kernels).  Every operation repeats the
// forward's (swe_device.h partwet1_w, swe_kernels.hip recon2_cell) in the
// same order, so the value parts are the forward's bit for bit; decisions
// are taken on the values.  The cube root and the bisection root carry
// their exact derivatives: d cbrt(y) = dy / (3 cbrt(y)^2), and the root x of
// x^3 + b x^2 + c x + d by the implicit function theorem, dx = -(db x^2 +
// dc x + dd) / (3 x^2 + 2 b x + c); an end of the bracket returned when the
// signs at the ends agree carries that end's derivative.
#pragma once

#include "swe_device.h"

namespace swe {

template <class T>
__device__ __forceinline__ T lift(double v) {
  T r = Zero<T>::get();
  r = r + v;
  return r;
}
template <>
__device__ __forceinline__ double lift<double>(double v) { return v; }

__device__ __forceinline__ double cbrt_g(double y) { return cbrt_det(y); }
template <int N>
__device__ __forceinline__ Dual<N> cbrt_g(const Dual<N>& y) {
  Dual<N> r;
  r.v = cbrt_det(y.v);
  const double g = r.v != 0. ? 1. / (3. * r.v * r.v) : 0.;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = y.d[k] * g;
  return r;
}

__device__ __forceinline__ double bisection_g(double b, double c, double d, double xmin, double xmax) {
  return bisection(Cubic{b, c, d}, xmin, xmax);
}
template <int N>
__device__ inline Dual<N> bisection_g(const Dual<N>& b, const Dual<N>& c, const Dual<N>& d,
                                      const Dual<N>& xmin, const Dual<N>& xmax) {
  const Cubic f{b.v, c.v, d.v};
  const double x = bisection(f, xmin.v, xmax.v);
  const double fmn = f(xmin.v), fmx = f(xmax.v);
  if (signbit(fmn) == signbit(fmx)) return (fabs(fmn) < fabs(fmx)) ? xmin : xmax;
  Dual<N> r;
  r.v = x;
  const double dp = (3. * x + 2. * b.v) * x + c.v;
  const double ig = dp != 0. ? -1. / dp : 0.;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = ((b.d[k] * x + c.d[k]) * x + d.d[k]) * ig;
  return r;
}

// ReconstructPartWetCell1: the volume-matching level (partwet1_w)
template <class T>
__device__ inline T partwet1_g(const T& z0, const T& z1, const T& z2, const T& bcell, const T& w) {
  const T b13 = dmax(dmax(z0, z1), z2);
  const T b23 = dmin(dmin(z0, z1), z2);
  const T b12 = 3. * bcell - b23 - b13;
  const T b_delim = b12 + (1. / 3.) * (b13 - b12) * (b13 - b12) / (b13 - b23);
  const T h = w - bcell;
  if (val(w) >= val(b13)) return w;
  if (val(w) <= val(b_delim)) return b23 + cbrt_g(3. * h * (b13 - b23) * (b12 - b23));
  const T a = -3. * b13;
  const T b = 3. * (b12 * b13 + b13 * b23 - b12 * b23);
  const T c = (b13 - b23) * (3. * h * (b13 - b12) - b12 * (b12 + b23)) - b23 * b23 * b13;
  return bisection_g(a, b, c, b12, b13);
}

// a point whose coordinates may all carry derivatives (PartWet2 moves its
// plane points along the cell's edges)
template <class T>
struct PT {
  T x, y, z;
};

// Gradient (src/PointOperations.cpp:42-48), pivot decided on the values
template <class T>
__device__ inline void gradient_g(const PT<T>& p0, const PT<T>& p1, const PT<T>& p2, T& gx, T& gy) {
  T a00 = p1.x - p0.x, a01 = p1.y - p0.y, b0 = p1.z - p0.z;
  T a10 = p2.x - p0.x, a11 = p2.y - p0.y, b1 = p2.z - p0.z;
  if (fabs(val(a10)) > fabs(val(a00))) {
    T t;
    t = a00; a00 = a10; a10 = t;
    t = a01; a01 = a11; a11 = t;
    t = b0; b0 = b1; b1 = t;
  }
  const T l = a10 / a00;
  const T u11 = a11 - l * a01;
  const T y1 = b1 - l * b0;
  const T x1 = y1 / u11;
  const T x0 = (b0 - a01 * x1) / a00;
  gx = x0;
  gy = x1;
}

// ReconstructPartWetCell2 after the fallback test: Q0, Q1, Q2 the vertices
// sorted by height (the forward's three conditional swaps), (cx, cy) the
// centroid, bc its height, w the level, w23 the pass-1 node maximum at Q0.
// Returns false where the forward falls back to the pass-1 level (the
// bisection's k1 below the tolerance).
template <class T>
__device__ inline bool partwet2_g(const PT<T>& Q0, const PT<T>& Q1, const PT<T>& Q2, double cx, double cy,
                                  const T& bc, const T& w, const T& w23, T& w0, T& gx, T& gy) {
  const T b23 = Q0.z, b12 = Q1.z, b13 = Q2.z;
  const T h23 = w23 - b23;
  const T ratio_b = (b12 - b23) / (b13 - b23);
  const T hd1 = 1. / 3. * h23 * ratio_b;
  const T hd2 = 1. / 3. * h23 * (2. * b13 - b12 - b23) / (b13 - b23);
  const T hi = w - bc;
  const T ratio_h = hi / h23;
  PT<T> p0 = Q0, p1, p2;
  p0.z = w23;
  if (val(hi) <= val(hd1)) {
    const T k2 = sqrt(3. * ratio_h / ratio_b);
    p1 = PT<T>{k2 * Q1.x + (1. - k2) * Q0.x, k2 * Q1.y + (1. - k2) * Q0.y, k2 * Q1.z + (1. - k2) * Q0.z};
    const T k3 = sqrt(3. * ratio_h * ratio_b);
    p2 = PT<T>{k3 * Q2.x + (1. - k3) * Q0.x, k3 * Q2.y + (1. - k3) * Q0.y, k3 * Q2.z + (1. - k3) * Q0.z};
  } else if (val(hi) >= val(hd2)) {
    const T delta_w = 1.5 * (hi - hd2);
    p1 = Q1;
    p1.z = p1.z + delta_w;
    p1.z = p1.z + (1. - ratio_b) * h23;
    p2 = Q2;
    p2.z = p2.z + delta_w;
  } else {
    const T alpha = 3. * ratio_h;
    const T beta = (b13 - b12) / (b13 - b23);
    const T k1 = 1. - bisection_g(Zero<T>::get(), (alpha - 3.) / beta, (1. + beta - alpha) / (beta * beta),
                                  Zero<T>::get(), lift<T>(1.));
    if (val(k1) < kTol) return false;
    const T k3 = 1. - beta * (1. - k1);
    const T bp1 = k1 * b13 + (1. - k1) * b12;
    p1 = Q1;
    p1.z = b12 + (k1 / k3) * beta * h23;
    p2 = PT<T>{k1 * Q2.x + (1. - k1) * Q1.x, k1 * Q2.y + (1. - k1) * Q1.y, k1 * Q2.z + (1. - k1) * Q1.z};
    p2.z = bp1;
  }
  gradient_g(p0, p1, p2, gx, gy);
  w0 = w23 + (gx * (cx - Q0.x) + gy * (cy - Q0.y));
  return true;
}

// sort three vertices by height with the forward's conditional swaps,
// carrying their node indices
template <class T>
__device__ __forceinline__ void sort3_g(PT<T>& Q0, PT<T>& Q1, PT<T>& Q2, int& i0, int& i1, int& i2) {
  PT<T> t;
  int ti;
  if (val(Q0.z) > val(Q1.z)) { t = Q0; Q0 = Q1; Q1 = t; ti = i0; i0 = i1; i1 = ti; }
  if (val(Q1.z) > val(Q2.z)) { t = Q1; Q1 = Q2; Q2 = t; ti = i1; i1 = i2; i2 = ti; }
  if (val(Q0.z) > val(Q1.z)) { t = Q0; Q0 = Q1; Q1 = t; ti = i0; i0 = i1; i1 = ti; }
}

}  // namespace swe
