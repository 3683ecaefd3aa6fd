// Device-side numerics of the explicit step: geometry, wet/dry MUSCL
// reconstruction, assigners, Riemann fluxes.  Each function restates the
// reference function it cites; expressions keep the reference's evaluation
// order and the library is built with -ffp-contract=off, so results are
// reproducible bit for bit against any implementation that follows the
// same steps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swe_dual.h"

namespace swe {

constexpr double kTol = 1e-13;  // include/Includes.h:30

struct V3 {
  double x, y, z;
};

// std::min / std::max semantics.
__host__ __device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// IsWet: include/Bathymetry.h:5-8
__host__ __device__ __forceinline__ bool is_wet(double h) { return h > 1e-12; }

// Cube root by a fixed sequence of IEEE operations: x = m 2^(3q), m in
// [0.5, 4), linear seed, five Newton steps y -= (y^3 - m) / (3 y^2).
// Replaces libm cbrt at src/MUSCLObject.cpp:103.
__device__ inline double cbrt_det(double x) {
  if (x == 0.0 || !isfinite(x)) return x;
  double s = 1.0;
  if (x < 0.0) { s = -1.0; x = -x; }
  int e;
  double m = frexp(x, &e);
  int r = e % 3;
  if (r < 0) r += 3;
  int q = (e - r) / 3;
  m = ldexp(m, r);
  double y = 0.7937005259840998 + (m - 0.5) * 0.22679491924311227;
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    double y2 = y * y;
    y = y - (y2 * y - m) / (3.0 * y2);
  }
  return s * ldexp(y, q);
}

// (int)log2(d), d > 0, from the exponent (src/PointOperations.cpp:33).
__device__ inline int ilog2_trunc(double d) {
  int e;
  double m = frexp(d, &e);
  int fl = e - 1;
  if (d >= 1.0) return fl;
  return (m == 0.5) ? fl : fl + 1;
}

// CubicPoly x^3 + b x^2 + c x + d (include/CubicPolyMath.h:6-19)
struct Cubic {
  double b, c, d;
  __device__ __forceinline__ double operator()(double x) const {
    return x * x * x + b * x * x + c * x + d;
  }
};

// Bisection: src/PointOperations.cpp:26-40
__device__ inline double bisection(const Cubic& f, double xmin, double xmax) {
  double fmn = f(xmin), fmx = f(xmax);
  if (signbit(fmn) == signbit(fmx)) return (fabs(fmn) < fabs(fmx)) ? xmin : xmax;
  double d = xmax - xmin;
  int n = (d > 0.0) ? 50 + ilog2_trunc(d) : -1;
  double x = 0.5 * (xmin + xmax);
  for (int i = 0; i <= n; i++) {
    x = 0.5 * (xmin + xmax);
    double fx = f(x);
    if (signbit(fmn) != signbit(fx)) {
      xmax = x;
    } else {
      xmin = x;
      fmn = fx;
    }
  }
  return x;
}

// Gradient: src/PointOperations.cpp:42-48 — plane through three points,
// 2x2 partial-pivot LU.
__device__ inline void gradient(const V3& p0, const V3& p1, const V3& p2, double& gx, double& gy) {
  double a00 = p1.x - p0.x, a01 = p1.y - p0.y, b0 = p1.z - p0.z;
  double a10 = p2.x - p0.x, a11 = p2.y - p0.y, b1 = p2.z - p0.z;
  if (fabs(a10) > fabs(a00)) {
    double t;
    t = a00; a00 = a10; a10 = t;
    t = a01; a01 = a11; a11 = t;
    t = b0; b0 = b1; b1 = t;
  }
  double l = a10 / a00;
  double u11 = a11 - l * a01;
  double y1 = b1 - l * b0;
  double x1 = y1 / u11;
  double x0 = (b0 - a01 * x1) / a00;
  gx = x0;
  gy = x1;
}

// Domain::T (src/Bathymetry.cpp:24-27): ((p0 c + p1 c) + p2 c), c = 1/3.
__device__ __forceinline__ V3 centroid(const V3& a, const V3& b, const V3& d) {
  const double c = 1. / 3.;
  return {(a.x * c + b.x * c) + d.x * c, (a.y * c + b.y * c) + d.y * c,
          (a.z * c + b.z * c) + d.z * c};
}
// Domain::E: midpoint of the two edge nodes.
__device__ __forceinline__ V3 midpoint(const V3& a, const V3& b) {
  return {0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)};
}
// Len (src/PointOperations.cpp:4-6)
__device__ __forceinline__ double edge_len(const V3& a, const V3& b) {
  return sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y));
}
// TriangArea (src/PointOperations.cpp:12-14)
__device__ __forceinline__ double tri_area(const V3& a, const V3& b, const V3& c) {
  double ux = b.x - a.x, uy = b.y - a.y, vx = c.x - a.x, vy = c.y - a.y;
  return 0.5 * fabs(ux * vy - vx * uy);
}
// Domain::Tang (src/Bathymetry.cpp:69-76): unit (P1 - P0) / L flipped to
// point outward of the cell with centroid c.  a = P(ep0), b = P(ep1).
__device__ __forceinline__ void tangent(const V3& a, const V3& b, double l, const V3& c,
                                        double& tx, double& ty) {
  tx = (b.x - a.x) / l;
  ty = (b.y - a.y) / l;
  double dx = c.x - a.x, dy = c.y - a.y;
  if (dx * ty - tx * dy > 0.) { tx = -tx; ty = -ty; }
}

// MUSCL::AtPoint (include/MUSCLObject.h:26-33), w component and wet test;
// velocity is the origin's (gradients of u, v are always zero).
__device__ __forceinline__ double at_point_w(double w0, double gx, double gy, const V3& c,
                                             const V3& p, bool& wet) {
  double dx = p.x - c.x, dy = p.y - c.y;
  double a0 = w0 + (gx * dx + gy * dy);
  wet = (a0 - p.z) >= 0;
  return wet ? a0 : p.z;
}

// PrimAssigner::operator= (src/Assigners.cpp:8-20) applied to AtPoint.
__device__ __forceinline__ void prim_assign(double b, double w, double u, double v, double& ow,
                                            double& ou, double& ov) {
  double h = w - b;
  if (!is_wet(h)) { ow = b; ou = 0.; ov = 0.; return; }
  ow = w; ou = u; ov = v;
  if (h < 1e-3) {
    double fac = sqrt(2.) * h / sqrt(h * h + 1e-6);
    ou *= fac;
    ov *= fac;
  }
}

// ConsAssigner::operator= (src/Assigners.cpp:28-44)
__device__ __forceinline__ void cons_assign(double b, double h, double hu, double hv, double& ow,
                                            double& ou, double& ov) {
  if (!is_wet(h)) { ow = b; ou = 0.; ov = 0.; return; }
  double ih;
  if (h < 1e-3) ih = sqrt(2.) * h / sqrt(h * h * h * h + 1e-12);
  else ih = 1. / h;
  ow = h + b;
  ou = hu * ih;
  ov = hv * ih;
}

// The flux functions are templates on the scalar type T of the states
// (double in the step; Dual<N> in the tangent-linear kernels, for the
// derivative of the very same operation sequence); geometry stays double.
// They are host-callable as well, for the host check of their reverse mode
// (swe_flux_adj.h) against the dual-number derivative.

// ElemFlux (include/SpaceDisc.h:4-18)
template <class T>
__host__ __device__ __forceinline__ void elem_flux(double nx, double ny, const T& h, const T& hu, const T& hv,
                                          T& f0, T& f1, T& f2) {
  f0 = Zero<T>::get(); f1 = f0; f2 = f0;
  if (is_wet(h)) {
    T hveln = hu * nx + hv * ny;
    f0 = hveln;
    f1 = (hveln / h) * hu + (0.5 * h * h) * nx;
    f2 = (hveln / h) * hv + (0.5 * h * h) * ny;
  }
}

// Wavespeeds (src/Fluxes.cpp:5-26)
template <int WS, class T>
__host__ __device__ __forceinline__ void wavespeed(const T& ul, const T& hl, const T& ur, const T& hr, T& a0,
                                          T& a1) {
  T cl = sqrt(hl), cr = sqrt(hr);
  if (WS == 0) {  // Rusanov
    T ap = dmax(fabs(ul) + cl, fabs(ur) + cr);
    a0 = -ap; a1 = ap;
  } else if (WS == 1) {  // Davis
    a0 = dmin(ul - cl, ur - cr); a1 = dmax(ul + cl, ur + cr);
  } else {  // Einfeldt, Roe velocity as in the reference
    T uRoe = (cl * ul + cl * ur) / (cl + cr);
    T cRoe = sqrt(0.5 * (hl + hr));
    a0 = dmin(ul - cl, uRoe - cRoe); a1 = dmax(ur + cr, uRoe + cRoe);
  }
}

// Edge state seen by the flux: primitive slot (w, u, v) and edge bathymetry.
template <class T>
struct SlotT {
  T w, u, v;
};
using Slot = SlotT<double>;

// Fluxes::HLL<W> (include/Fluxes.h:14-54).  lw: length-to-wavespeed
// candidate for the CFL minimum (unchanged when the flux returns early);
// dmn = min(2 Area(from) / L, 2 Area(to) / L).  eb (SWE_OPT_EXACT_BALANCE):
// the same flux written Fl + al (ar (Ur - Ul) - (Fr - Fl)) / (ar - al),
// which is Fl to the bit when the two states are equal.
template <int WS, class T>
__host__ __device__ __forceinline__ void flux_hll(double nx, double ny, const T& be, const SlotT<T>& L,
                                         const SlotT<T>& R, double dmn, double cor, double& lw,
                                         T F[3], bool eb = false) {
  T ul = L.u * nx + L.v * ny;
  T hl = L.w - be;
  T ur = R.u * nx + R.v * ny;
  T hr = R.w - be;
  F[0] = Zero<T>::get(); F[1] = F[0]; F[2] = F[0];
  if (hl + hr <= 1e-10) return;
  T a0, a1;
  wavespeed<WS>(ul, hl, ur, hr, a0, a1);
  T al = dmin(0., a0), ar = dmax(0., a1);
  if (ar - al <= 1e-10) return;
  double c = fabs(cor);
  lw = dmn / (c + dmax(-val(al), val(ar)));
  T Ul0 = hl, Ul1 = hl * L.u, Ul2 = hl * L.v;
  T Ur0 = hr, Ur1 = hr * R.u, Ur2 = hr * R.v;
  T Fl0, Fl1, Fl2, Fr0, Fr1, Fr2;
  elem_flux(nx, ny, Ul0, Ul1, Ul2, Fl0, Fl1, Fl2);
  elem_flux(nx, ny, Ur0, Ur1, Ur2, Fr0, Fr1, Fr2);
  T aa = al * ar, den = ar - al;
  if (eb) {
    F[0] = Fl0 + al * (ar * (Ur0 - Ul0) - (Fr0 - Fl0)) / den;
    F[1] = Fl1 + al * (ar * (Ur1 - Ul1) - (Fr1 - Fl1)) / den;
    F[2] = Fl2 + al * (ar * (Ur2 - Ul2) - (Fr2 - Fl2)) / den;
    return;
  }
  F[0] = (ar * Fl0 - al * Fr0 + aa * (Ur0 - Ul0)) / den;
  F[1] = (ar * Fl1 - al * Fr1 + aa * (Ur1 - Ul1)) / den;
  F[2] = (ar * Fl2 - al * Fr2 + aa * (Ur2 - Ul2)) / den;
}

// Fluxes::HLLC<W> (include/Fluxes.h:56-111).  eb (SWE_OPT_EXACT_BALANCE):
// the star depth takes the speed ratio first, h (a - u) / (a - u*) as
// h ((a - u) / (a - u*)), so that at rest (u = u* = 0) it is h exactly.
template <int WS, class T>
__host__ __device__ __forceinline__ void flux_hllc(double tx, double ty, const T& be, const SlotT<T>& L,
                                          const SlotT<T>& R, double dmn, double cor, double& lw,
                                          T F[3], bool eb = false) {
  double nx = ty, ny = -tx;
  T ul = L.u * nx + L.v * ny;
  T vl = L.u * tx + L.v * ty;
  T hl = L.w - be;
  T ur = R.u * nx + R.v * ny;
  T vr = R.u * tx + R.v * ty;
  T hr = R.w - be;
  F[0] = Zero<T>::get(); F[1] = F[0]; F[2] = F[0];
  if (hl + hr <= 1e-10) return;
  T al, ar;
  wavespeed<WS>(ul, hl, ur, hr, al, ar);
  T ustar = (ar - ur) * hr * ur - (al - ul) * hl * ul + 0.5 * (hl * hl - hr * hr);
  ustar = ustar / (hr * (ar - ur) - hl * (al - ul));
  double c = fabs(cor);
  lw = dmn / (c + dmax(kTol, dmax(val(al), val(ar))));
  T U0, U1, U2, Fe0, Fe1, Fe2;
  if (ustar <= 0) {
    T us = vr * tx + ustar * ty;
    T vs = vr * nx + ustar * ny;
    T hs;
    if (eb) hs = hr * ((ar - ur) / (ar - ustar));
    else hs = hr * (ar - ur) / (ar - ustar);
    U0 = hr; U1 = hr * R.u; U2 = hr * R.v;
    elem_flux(nx, ny, U0, U1, U2, Fe0, Fe1, Fe2);
    T s = dmax(0., ar);
    F[0] = Fe0 + s * (hs - U0);
    F[1] = Fe1 + s * (hs * us - U1);
    F[2] = Fe2 + s * (hs * vs - U2);
  } else {
    T us = vl * tx + ustar * ty;
    T vs = vl * nx + ustar * ny;
    T hs;
    if (eb) hs = hl * ((al - ul) / (al - ustar));
    else hs = hl * (al - ul) / (al - ustar);
    U0 = hl; U1 = hl * L.u; U2 = hl * L.v;
    elem_flux(nx, ny, U0, U1, U2, Fe0, Fe1, Fe2);
    T s = dmin(0., al);
    F[0] = Fe0 + s * (hs - U0);
    F[1] = Fe1 + s * (hs * us - U1);
    F[2] = Fe2 + s * (hs * vs - U2);
  }
}

// ReconstructPartWetCell1 (src/MUSCLObject.cpp:86-112): volume-matching
// constant level of a partially wet cell; returns the reconstructed w.
__device__ inline double partwet1_w(double z0, double z1, double z2, double bcell, double w) {
  double b13 = dmax(dmax(z0, z1), z2);
  double b23 = dmin(dmin(z0, z1), z2);
  double b12 = 3. * bcell - b23 - b13;
  double b_delim = b12 + (1. / 3.) * (b13 - b12) * (b13 - b12) / (b13 - b23);
  double h = w - bcell;
  if (w >= b13) return w;
  if (w <= b_delim) return b23 + cbrt_det(3. * h * (b13 - b23) * (b12 - b23));
  double a = -3. * b13;
  double b = 3. * (b12 * b13 + b13 * b23 - b12 * b23);
  double c = (b13 - b23) * (3. * h * (b13 - b12) - b12 * (b12 + b23)) - b23 * b23 * b13;
  return bisection(Cubic{a, b, c}, b12, b13);
}

}  // namespace swe
