// Tangent-linear step: the forward stage kernels built on dual numbers.
//
// Each kernel repeats the operation sequence of its forward counterpart
// (k_recon1, k_recon2, k_flux, k_drain, k_update with
// TimeDisc::RHS, src/SpaceDisc.cpp:15-74, src/TimeDisc.cpp:3-66,
// src/Assigners.cpp:8-44) with the state, the node bathymetry and every
// quantity derived from them carried as Dual<1>: the value parts are the
// forward's, bit for bit, and the derivative parts are the directional
// derivative of the step along (state tangent, bathymetry tangent).  Branch
// decisions are taken on the values, as the forward takes them.  Geometry in
// the plane (x, y) has no tangent.
//
// Partially wet cells: PartWet1's level (cube root or bisection root, with
// their implicit derivatives) in pass 1, PartWet2's plane in pass 2
// (k_tl_recon2) through the pass-1 node maximum, whose tangent is that of
// the cell attaining it (swe_pw.h).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "swe_common.h"
#include "swe_pw.h"

namespace swe {

namespace {

using D = Dual<1>;

__device__ __forceinline__ D dd(double v, double d) {
  D r;
  r.v = v;
  r.d[0] = d;
  return r;
}
__device__ __forceinline__ D dzero() { return dd(0., 0.); }

// a point whose height carries a tangent
struct P3 {
  double x, y;
  D z;
};
__device__ __forceinline__ P3 node_t(const DevMesh& M, const DevTan& Q, int n) {
  return {M.px[n], M.py[n], dd(M.pz[n], Q.zd[n])};
}
__device__ __forceinline__ P3 pick_t(const P3& a, const P3& b, const P3& c, int i) {
  return i == 0 ? a : (i == 1 ? b : c);
}
// Domain::T and Domain::E in the forward's operation order
__device__ __forceinline__ P3 centroid_t(const P3& a, const P3& b, const P3& d) {
  const double c = 1. / 3.;
  return {(a.x * c + b.x * c) + d.x * c, (a.y * c + b.y * c) + d.y * c, (a.z * c + b.z * c) + d.z * c};
}
__device__ __forceinline__ P3 midpoint_t(const P3& a, const P3& b) {
  return {0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)};
}
__device__ __forceinline__ V3 flat(const P3& p) { return {p.x, p.y, 0.}; }

// Gradient (src/PointOperations.cpp:42-48) with tangent-carrying heights
__device__ inline void gradient_t(const P3& p0, const P3& p1, const P3& p2, D& gx, D& gy) {
  double a00 = p1.x - p0.x, a01 = p1.y - p0.y;
  D b0 = p1.z - p0.z;
  double a10 = p2.x - p0.x, a11 = p2.y - p0.y;
  D b1 = p2.z - p0.z;
  if (fabs(a10) > fabs(a00)) {
    double t;
    t = a00; a00 = a10; a10 = t;
    t = a01; a01 = a11; a11 = t;
    const D tb = b0; b0 = b1; b1 = tb;
  }
  const double l = a10 / a00;
  const double u11 = a11 - l * a01;
  const D y1 = b1 - l * b0;
  const D x1 = y1 / u11;
  const D x0 = (b0 - a01 * x1) / a00;
  gx = x0;
  gy = x1;
}

struct MusclT {
  D w0, gx, gy, u0, v0;
};
__device__ __forceinline__ MusclT load_muscl_t(const DevState& S, const DevState& Sd, const DevWork& W,
                                               const DevTan& Q, int c) {
  MusclT m{dd(W.mw[c], Q.mw[c]), dd(W.mgx[c], Q.mgx[c]), dd(W.mgy[c], Q.mgy[c]), dd(S.u[c], Sd.u[c]),
           dd(S.v[c], Sd.v[c])};
  if (W.cls[c] == 0) {
    m.u0 = dzero();
    m.v0 = dzero();
  }
  return m;
}

// PrimAssigner over AtPoint (slot_value)
__device__ __forceinline__ void prim_assign_t(const D& b, const D& w, const D& u, const D& v, D& ow, D& ou,
                                              D& ov) {
  const D h = w - b;
  if (!is_wet(h)) {
    ow = b; ou = dzero(); ov = dzero();
    return;
  }
  ow = w; ou = u; ov = v;
  if (h < 1e-3) {
    const D fac = sqrt(2.) * h / sqrt(h * h + 1e-6);
    ou = ou * fac;
    ov = ov * fac;
  }
}
__device__ __forceinline__ SlotT<D> slot_value_t(const MusclT& m, double cx, double cy, const P3& E) {
  const double dx = E.x - cx, dy = E.y - cy;
  D aw = m.w0 + (m.gx * dx + m.gy * dy);
  const bool wet = (aw - E.z) >= 0;
  SlotT<D> s;
  if (!wet) aw = E.z;
  prim_assign_t(E.z, aw, wet ? m.u0 : dzero(), wet ? m.v0 : dzero(), s.w, s.u, s.v);
  return s;
}

}  // namespace

// ------------------------------------------------------ reconstruction --
// pass 1 (k_recon1): dry, fully wet and the partially wet cells' PartWet1
// level (kept in wrec1 for the node maxima of pass 2)
__global__ void __launch_bounds__(256) k_tl_recon(DevMesh M, DevState S, DevState Sd, DevWork W, DevTan Q) {
  const int64_t i64 = global_index();  // XCD-aware, as the forward kernels
  if (i64 >= M.T) return;
  const int i = (int)i64;
  const int64_t T = M.T;
  const D w = dd(S.w[i], Sd.w[i]);
  const uint32_t fl = M.cflag[i];
  const P3 p0 = node_t(M, Q, M.tp[i]), p1 = node_t(M, Q, M.tp[T + i]), p2 = node_t(M, Q, M.tp[2 * T + i]);
  const D bc = centroid_t(p0, p1, p2).z;
  const CellRec ri = M.crec[i];
  D w0 = w, gx = dzero(), gy = dzero();
  int8_t k;
  if (!is_wet(w - bc)) {
    k = 0;
    w0 = bc;
    gradient_t(p0, p1, p2, gx, gy);
  } else if ((fl & CELL_BND) || !(ri.zmax < w.v)) {
    k = 1;
    const D zmn = dmin(dmin(p0.z, p1.z), p2.z), zmx = dmax(dmax(p0.z, p1.z), p2.z);
    w0 = partwet1_g(zmn, zmn, zmx, bc, w);
    W.wrec1[i] = w0.v;
    Q.wrec1[i] = w0.d[0];
  } else {
    k = 2;
    P3 pts[3], E[3];
    bool any_dry = false;
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      E[kk] = midpoint_t(pick_t(p0, p1, p2, ep0_local(fl, kk)), pick_t(p0, p1, p2, ep1_local(fl, kk)));
      const int j = M.tt[kk * T + i];  // full-wet cells are interior
      const CellRec rj = M.crec[j];
      const D wj = dd(S.w[j], Sd.w[j]);
      if (!(rj.fl & CELL_BND) && rj.zmax < wj.v) {
        pts[kk] = {rj.cx, rj.cy, wj};
      } else if (!is_wet(wj.v - rj.bc)) {
        any_dry = true;
        pts[kk] = {0., 0., dzero()};
      } else {
        // PartWet1 level of the partially wet neighbour, from its vertices
        const P3 q0 = node_t(M, Q, M.tp[j]), q1 = node_t(M, Q, M.tp[T + j]), q2 = node_t(M, Q, M.tp[2 * T + j]);
        const D zmnj = dmin(dmin(q0.z, q1.z), q2.z), zmxj = dmax(dmax(q0.z, q1.z), q2.z);
        const D w0j = partwet1_g(zmnj, zmnj, zmxj, centroid_t(q0, q1, q2).z, wj);
        const double dx = E[kk].x - rj.cx, dy = E[kk].y - rj.cy;
        D a = w0j + (0. * dx + 0. * dy);
        if (!((a - E[kk].z) >= 0)) a = E[kk].z;
        pts[kk] = {E[kk].x, E[kk].y, 0.5 * (w + a)};
      }
    }
    if (!any_dry) {
      gradient_t(pts[0], pts[1], pts[2], gx, gy);
      const double c23 = 1. - 1. / 3., c13 = -1. / 3.;
      bool all_wet = true;
      {
        double dxm = (p0.x * c23 + p1.x * c13) + p2.x * c13, dym = (p0.y * c23 + p1.y * c13) + p2.y * c13;
        if (!is_wet(((gx.v * dxm + gy.v * dym) + w.v) - p0.z.v)) all_wet = false;
      }
      {
        double dxm = (p0.x * c13 + p1.x * c23) + p2.x * c13, dym = (p0.y * c13 + p1.y * c23) + p2.y * c13;
        if (!is_wet(((gx.v * dxm + gy.v * dym) + w.v) - p1.z.v)) all_wet = false;
      }
      {
        double dxm = (p0.x * c13 + p1.x * c13) + p2.x * c23, dym = (p0.y * c13 + p1.y * c13) + p2.y * c23;
        if (!is_wet(((gx.v * dxm + gy.v * dym) + w.v) - p2.z.v)) all_wet = false;
      }
      if (!all_wet) { gx = dzero(); gy = dzero(); }
      double tvd = 1.;
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        const double wn = S.w[M.tt[kk * T + i]];
        const double vtmin = dmin(w.v, wn), vtmax = dmax(w.v, wn);
        const double vek = w.v + (gx.v * (E[kk].x - ri.cx) + gy.v * (E[kk].y - ri.cy));
        if (!((vtmin <= vek) && (vek <= vtmax))) tvd = 0.;
      }
      gx = tvd * gx;
      gy = tvd * gy;
    }
  }
  W.mw[i] = w0.v;
  W.mgx[i] = gx.v;
  W.mgy[i] = gy.v;
  W.cls[i] = k;
  Q.mw[i] = w0.d[0];
  Q.mgx[i] = gx.d[0];
  Q.mgy[i] = gy.d[0];
}

// pass 2 (recon2_cell): PartWet2 of the partially wet cells through the
// pass-1 node maximum at their lowest vertex (node_max_pass1), whose tangent
// is the attaining cell's pass-1 tangent (or the node's own height)
__device__ inline D node_max_t(const DevMesh& M, const DevWork& W, const DevTan& Q, int n, const P3& P) {
  D mx = P.z;
  for (int q = M.n2c_off[n]; q < M.n2c_off[n + 1]; ++q) {
    const int c = M.n2c[q];
    D w0 = dd(W.mw[c], Q.mw[c]), gx = dd(W.mgx[c], Q.mgx[c]), gy = dd(W.mgy[c], Q.mgy[c]);
    if (W.cls[c] == 1) {
      w0 = dd(W.wrec1[c], Q.wrec1[c]);
      gx = dzero();
      gy = dzero();
    }
    const double dx = P.x - M.cx[c], dy = P.y - M.cy[c];
    D a = w0 + (gx * dx + gy * dy);
    if (!((a - P.z) >= 0)) a = P.z;
    mx = dmax(mx, a);
  }
  return mx;
}

__global__ void __launch_bounds__(256) k_tl_recon2(DevMesh M, DevState S, DevState Sd, DevWork W, DevTan Q) {
  const int64_t i64 = global_index();  // XCD-aware, as the forward kernels
  if (i64 >= M.T) return;
  const int i = (int)i64;
  if (W.cls[i] != 1) return;
  const int64_t T = M.T;
  int i0 = M.tp[i], i1 = M.tp[T + i], i2 = M.tp[2 * T + i];
  const P3 v0 = node_t(M, Q, i0), v1 = node_t(M, Q, i1), v2 = node_t(M, Q, i2);
  const D bc = centroid_t(v0, v1, v2).z;
  PT<D> Q0{dd(v0.x, 0.), dd(v0.y, 0.), v0.z}, Q1{dd(v1.x, 0.), dd(v1.y, 0.), v1.z}, Q2{dd(v2.x, 0.), dd(v2.y, 0.), v2.z};
  sort3_g(Q0, Q1, Q2, i0, i1, i2);
  const D w = dd(S.w[i], Sd.w[i]);
  D w0 = dd(W.wrec1[i], Q.wrec1[i]), gx = dzero(), gy = dzero();  // fallback: the pass-1 level
  const bool fallback = (w.v > Q2.z.v) || (Q2.z.v - Q0.z.v < kTol);
  if (!fallback) {
    const D w23 = node_max_t(M, W, Q, i0, P3{Q0.x.v, Q0.y.v, Q0.z});
    D a, bx, by;
    if (partwet2_g(Q0, Q1, Q2, M.cx[i], M.cy[i], bc, w, w23, a, bx, by)) {
      w0 = a;
      gx = bx;
      gy = by;
    }
  }
  W.mw[i] = w0.v;
  W.mgx[i] = gx.v;
  W.mgy[i] = gy.v;
  Q.mw[i] = w0.d[0];
  Q.mgx[i] = gx.d[0];
  Q.mgy[i] = gy.d[0];
}

// ----------------------------------------------------------------- flux --
template <int FLUX, int WS>
__global__ void __launch_bounds__(256) k_tl_flux(DevMesh M, DevState S, DevState Sd, DevWork W, DevTan Q,
                                                 double cor) {
  const int64_t e64 = global_index();  // XCD-aware, as the forward kernels
  if (e64 >= M.E) return;
  const int e = (int)e64;
  const int64_t T = M.T;
  const int f = M.et0[e], t = M.et1[e];
  const P3 Pa = node_t(M, Q, M.ep0[e]), Pb = node_t(M, Q, M.ep1[e]);
  const double le = edge_len(flat(Pa), flat(Pb));
  const P3 E = midpoint_t(Pa, Pb);
  const double cxf = M.cx[f], cyf = M.cy[f];
  double tx, ty;
  tangent(flat(Pa), flat(Pb), le, V3{cxf, cyf, 0.}, tx, ty);
  const double nx = ty, ny = -tx;
  D F[3];
  double lw = 1.;
  if (t == kSolidWall) {
    const D bcf = centroid_t(node_t(M, Q, M.tp[f]), node_t(M, Q, M.tp[T + f]), node_t(M, Q, M.tp[2 * T + f])).z;
    const D hf = dd(S.w[f], Sd.w[f]) - bcf;
    elem_flux(nx, ny, hf, dzero(), dzero(), F[0], F[1], F[2]);
  } else {
    const int tr = t < 0 ? f : t;
    const SlotT<D> L = slot_value_t(load_muscl_t(S, Sd, W, Q, f), cxf, cyf, E);
    const SlotT<D> R = slot_value_t(load_muscl_t(S, Sd, W, Q, tr), M.cx[tr], M.cy[tr], E);
    if (FLUX == 0) flux_hll<WS>(nx, ny, E.z, L, R, M.edmin[e], cor, lw, F);
    else flux_hllc<WS>(tx, ty, E.z, L, R, M.edmin[e], cor, lw, F);
  }
  W.f0[e] = F[0].v;
  W.f1[e] = F[1].v;
  W.f2[e] = F[2].v;
  Q.f0[e] = F[0].d[0];
  Q.f1[e] = F[1].d[0];
  Q.f2[e] = F[2].d[0];
}

// ---------------------------------------------------------- draining dt --
template <bool LEN>
__global__ void __launch_bounds__(256) k_tl_drain(DevMesh M, DevState S, DevState Sd, DevWork W, DevTan Q) {
  const int64_t i64 = global_index();  // XCD-aware, as the forward kernels
  if (i64 >= M.T) return;
  const int i = (int)i64;
  const int64_t T = M.T;
  const P3 p0 = node_t(M, Q, M.tp[i]), p1 = node_t(M, Q, M.tp[T + i]), p2 = node_t(M, Q, M.tp[2 * T + i]);
  const D h = dd(S.w[i], Sd.w[i]) - centroid_t(p0, p1, p2).z;
  D d;
  if (!is_wet(h)) {
    d = dzero();
  } else {
    const uint32_t fl = M.cflag[i];
    D sum = dzero();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int e = M.te[k * T + i];
      const D fe = dd(W.f0[e], Q.f0[e]);
      D out = sgn_pos(fl, k) ? fe : -fe;
      if (LEN)
        out = out * edge_len(flat(pick_t(p0, p1, p2, ep0_local(fl, k))), flat(pick_t(p0, p1, p2, ep1_local(fl, k))));
      sum = sum + dmax(0., out);
    }
    d = sum > kTol ? M.area[i] * h / sum : dd(INFINITY, 0.);
  }
  W.dtd[i] = d.v;
  Q.dtd[i] = d.d[0];
}

// --------------------------------------------------------------- update --
__global__ void __launch_bounds__(256) k_tl_update(DevMesh M, DevState S, DevState Sd, DevState U0,
                                                   DevState U0d, DevOut O, DevOut Od, DevWork W, DevTan Q,
                                                   StageArgs A) {
  const int64_t i64 = global_index();  // XCD-aware, as the forward kernels
  if (i64 >= M.T) return;
  const int i = (int)i64;
  const int64_t T = M.T;
  const uint32_t fl = M.cflag[i];
  const MusclT mu = load_muscl_t(S, Sd, W, Q, i);
  const P3 p[3] = {node_t(M, Q, M.tp[i]), node_t(M, Q, M.tp[T + i]), node_t(M, Q, M.tp[2 * T + i])};
  const D dti = dd(W.dtd[i], Q.dtd[i]);
  const double dts = A.coef * A.dt_fixed;
  // TimeDisc::RHS
  D R0 = dzero(), R1 = dzero(), R2 = dzero();
  const P3 c = centroid_t(p[0], p[1], p[2]);
  const double i_area = 1. / tri_area(flat(p[0]), flat(p[1]), flat(p[2]));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int e = M.te[k * T + i];
    const int nb = M.tt[k * T + i];
    const bool own = sgn_pos(fl, k);
    const double sg = own ? 1. : -1.;
    const D F0 = dd(W.f0[e], Q.f0[e]), F1 = dd(W.f1[e], Q.f1[e]), F2 = dd(W.f2[e], Q.f2[e]);
    const D dtik = nb < 0 ? dd(INFINITY, 0.) : dd(W.dtd[nb], Q.dtd[nb]);
    const D dtk = (sg * F0) > 0. ? dmin(dts, dti) : dmin(dts, dtik);
    const P3 Pa = pick_t(p[0], p[1], p[2], ep0_local(fl, k));
    const P3 Pb = pick_t(p[0], p[1], p[2], ep1_local(fl, k));
    const double le = edge_len(flat(Pa), flat(Pb));
    const double c_ek = i_area * le;
    const P3 E = midpoint_t(Pa, Pb);
    const SlotT<D> s = slot_value_t(mu, c.x, c.y, E);
    const D h_ek = s.w - E.z;
    D src1 = mu.gx, src2 = mu.gy;
    src1 = src1 + A.cor * (-s.v);
    src2 = src2 + A.cor * s.u;
    double tx, ty;
    tangent(flat(Pa), flat(Pb), le, V3{c.x, c.y, 0.}, tx, ty);
    const double nx = ty, ny = -tx;
    const D cf = dtk * sg * c_ek;
    R0 = R0 - cf * F0;
    R1 = R1 - cf * F1;
    R2 = R2 - cf * F2;
    const double cs = dts * (1. / 3.);
    R1 = R1 - cs * src1 * h_ek;
    R2 = R2 - cs * src2 * h_ek;
    const D q2 = 0.5 * h_ek * h_ek;
    R1 = R1 + dtk * (nx * c_ek * q2);
    R2 = R2 + dtk * (ny * c_ek * q2);
  }
  const D w = dd(S.w[i], Sd.w[i]), u = dd(S.u[i], Sd.u[i]), v = dd(S.v[i], Sd.v[i]);
  if (A.fric != 0.) {
    const D hh = w - c.z;
    R1 = R1 - dts * (A.fric * (hh * mu.u0));
    R2 = R2 - dts * (A.fric * (hh * mu.v0));
  }
  // stage combination and ConsAssigner
  const D bc = c.z;
  const D h = w - bc;
  const D Uc0 = h, Uc1 = h * u, Uc2 = h * v;
  D r0, r1, r2;
  if (A.a0 != 0.) {
    const D h0 = dd(U0.w[i], U0d.w[i]) - bc;
    const D Ua0 = h0, Ua1 = h0 * dd(U0.u[i], U0d.u[i]), Ua2 = h0 * dd(U0.v[i], U0d.v[i]);
    r0 = (A.a0 * Ua0 + A.a1 * Uc0) + R0;
    r1 = (A.a0 * Ua1 + A.a1 * Uc1) + R1;
    r2 = (A.a0 * Ua2 + A.a1 * Uc2) + R2;
  } else {
    r0 = Uc0 + R0;
    r1 = Uc1 + R1;
    r2 = Uc2 + R2;
  }
  D ow, ou, ov;
  if (!is_wet(r0)) {
    ow = bc; ou = dzero(); ov = dzero();
  } else {
    D ih;
    if (r0 < 1e-3) ih = sqrt(2.) * r0 / sqrt(r0 * r0 * r0 * r0 + 1e-12);
    else ih = 1. / r0;
    ow = r0 + bc;
    ou = r1 * ih;
    ov = r2 * ih;
  }
  O.w[i] = ow.v; O.u[i] = ou.v; O.v[i] = ov.v;
  Od.w[i] = ow.d[0]; Od.u[i] = ou.d[0]; Od.v[i] = ov.d[0];
}


hipError_t launch_tangent_stage(const DevMesh& M, const DevState& S, const DevState& Sd, const DevState& U0,
                                const DevState& U0d, const DevOut& O, const DevOut& Od, const DevWork& W,
                                const DevTan& Q, const StageArgs& A, int flux, int ws, int opts,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_tl_recon, grid_for(M.T), dim3(256), 0, st, M, S, Sd, W, Q);
  hipLaunchKernelGGL(k_tl_recon2, grid_for(M.T), dim3(256), 0, st, M, S, Sd, W, Q);
#define SWE_TL_FLUX(F, Wv)                                                                         \
  if (flux == F && ws == Wv)                                                                       \
    hipLaunchKernelGGL((k_tl_flux<F, Wv>), grid_for(M.E), dim3(256), 0, st, M, S, Sd, W, Q, A.cor);
  SWE_TL_FLUX(0, 0) SWE_TL_FLUX(0, 1) SWE_TL_FLUX(0, 2)
  SWE_TL_FLUX(1, 0) SWE_TL_FLUX(1, 1) SWE_TL_FLUX(1, 2)
#undef SWE_TL_FLUX
  if (opts & kOptDrainLength) hipLaunchKernelGGL(k_tl_drain<true>, grid_for(M.T), dim3(256), 0, st, M, S, Sd, W, Q);
  else hipLaunchKernelGGL(k_tl_drain<false>, grid_for(M.T), dim3(256), 0, st, M, S, Sd, W, Q);
  hipLaunchKernelGGL(k_tl_update, grid_for(M.T), dim3(256), 0, st, M, S, Sd, U0, U0d, O, Od, W, Q, A);
  return hipGetLastError();
}

}  // namespace swe
