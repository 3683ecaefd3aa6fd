// gfx950 kernels of the explicit finite-volume step.
//
// Per stage (Solvers::Euler/SSPRK2/SSPRK3 stage, src/Solvers.cpp:5-59):
//   k_recon1  per cell  — classification + pass-1 MUSCL (dry / partially wet
//                         level / full-wet limited plane), stored as
//                         (w0, gx, gy) per cell: the reference's edge and
//                         source slots are functions of it and are never
//                         materialised on the hot path.
//   k_recon2  per cell  — pass 2 for partially wet cells (Liu et al. plane)
//                         with the node maximum gathered over the node->cell
//                         map (no atomics, order-independent max).
//   k_flux    per edge  — HLL/HLLC numerical flux from the two cells' MUSCL
//                         at the edge midpoint; wave-level + block min of the
//                         CFL length-to-wavespeed, one 64-bit atomic min per
//                         block on the positive-double bit pattern.
//   k_drain   per cell  — draining time step from the stage-start state.
//   k_update  per cell  — fixed-order gather of the three edge fluxes, the
//                         well-balanced source, SSP-RK stage combination and
//                         ConsAssigner write-back.
// Every reduction is either order-independent (max, min) or a fixed-order
// per-cell gather, so results are deterministic run to run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "swe_device.h"
#include "swe_launch.h"

namespace swe {

#define CELL_BND (1u << 3)

__device__ __forceinline__ int logical_block() {
  // Workgroups are dispatched round-robin over the 8 XCDs; hand each XCD a
  // contiguous range of the Morton-ordered cells so neighbour gathers hit
  // that XCD's L2.  gridDim.x is a multiple of 8.
  int per = gridDim.x >> 3;
  int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}
__device__ __forceinline__ int64_t global_index() {
  return (int64_t)logical_block() * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ V3 node(const DevMesh& M, int n) {
  return {M.px[n], M.py[n], M.pz[n]};
}
__device__ __forceinline__ void cell_nodes(const DevMesh& M, int i, V3 p[3]) {
  p[0] = node(M, M.tp[i]);
  p[1] = node(M, M.tp[M.T + i]);
  p[2] = node(M, M.tp[2 * M.T + i]);
}
__device__ __forceinline__ int ep0_local(uint32_t fl, int k) { return (fl >> (4 + 4 * k)) & 3; }
__device__ __forceinline__ int ep1_local(uint32_t fl, int k) { return (fl >> (6 + 4 * k)) & 3; }
__device__ __forceinline__ bool sgn_pos(uint32_t fl, int k) { return (fl >> k) & 1u; }

// Final MUSCL of cell c: pass-2 result for partially wet cells.
struct Muscl {
  double w0, gx, gy, u0, v0;
};
__device__ __forceinline__ Muscl load_muscl(const DevMesh& M, const DevState& S,
                                            const DevWork& W, int c) {
  int k = W.cls[c];
  Muscl m;
  if (k == 1) {
    m.w0 = W.aw[c]; m.gx = W.agx[c]; m.gy = W.agy[c];
  } else {
    m.w0 = W.mw[c]; m.gx = W.mgx[c]; m.gy = W.mgy[c];
  }
  if (k == 0) {
    m.u0 = 0.; m.v0 = 0.;
  } else {
    m.u0 = S.u[c]; m.v0 = S.v[c];
  }
  return m;
}

// AtPoint + PrimAssigner: the edge slot written by UpdateInterfaceValues
// (src/SpaceDisc.cpp:24-25).
__device__ __forceinline__ Slot slot_value(const Muscl& m, const V3& c, const V3& E) {
  bool wet;
  double aw = at_point_w(m.w0, m.gx, m.gy, c, E, wet);
  double au = wet ? m.u0 : 0., av = wet ? m.v0 : 0.;
  Slot s;
  prim_assign(E.z, aw, au, av, s.w, s.u, s.v);
  return s;
}

// ---------------------------------------------------------------- pass 1 --
// SpaceDisc::ComputeInterfaceValues pass 1 (src/SpaceDisc.cpp:37-45) with
// ReconstructDryCell / ReconstructPartWetCell1 / ReconstructFullWetCell
// (src/MUSCLObject.cpp:31-112).
template <bool HEAD_GRAD>
__global__ void __launch_bounds__(256) k_recon1(DevMesh M, DevState S, DevWork W) {
  int64_t i = global_index();
  if (i >= M.T) return;
  V3 p[3];
  cell_nodes(M, (int)i, p);
  V3 c = centroid(p[0], p[1], p[2]);
  double w = S.w[i];
  double bc = c.z;
  uint32_t fl = M.cflag[i];
  double w0 = w, gx = 0., gy = 0.;
  int8_t k;
  if (!is_wet(w - bc)) {
    k = 0;
    w0 = bc;
    gradient(p[0], p[1], p[2], gx, gy);
  } else if ((fl & CELL_BND) || !(dmax(dmax(p[0].z, p[1].z), p[2].z) < w)) {
    k = 1;
    w0 = partwet1_w(p[0].z, p[1].z, p[2].z, bc, w);
  } else {
    k = 2;
    V3 pts[3];
    V3 E[3];
    bool any_dry = false;
    double wn[3];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      E[kk] = midpoint(p[ep0_local(fl, kk)], p[ep1_local(fl, kk)]);
      int j = M.tt[kk * M.T + i];
      V3 q[3];
      cell_nodes(M, j, q);
      V3 cj = centroid(q[0], q[1], q[2]);
      double wj = S.w[j];
      wn[kk] = wj;
      uint32_t flj = M.cflag[j];
      if (!(flj & CELL_BND) && dmax(dmax(q[0].z, q[1].z), q[2].z) < wj) {
        pts[kk] = {cj.x, cj.y, HEAD_GRAD ? cj.z : wj};
      } else if (!is_wet(wj - cj.z)) {
        any_dry = true;
      } else {
        double w0j = partwet1_w(q[0].z, q[1].z, q[2].z, cj.z, wj);
        bool wet;
        double a = at_point_w(w0j, 0., 0., cj, E[kk], wet);
        double val = 0.5 * (w + a);
        pts[kk] = {E[kk].x, E[kk].y, HEAD_GRAD ? E[kk].z : val};
      }
      if (any_dry) break;
    }
    if (!any_dry) {
      gradient(pts[0], pts[1], pts[2], gx, gy);
      // vertex positivity (src/MUSCLObject.cpp:66-72)
      const double c23 = 1. - 1. / 3., c13 = -1. / 3.;
      bool all_wet = true;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        double wx0 = p[0].x * (m == 0 ? c23 : c13), wy0 = p[0].y * (m == 0 ? c23 : c13);
        double wx1 = p[1].x * (m == 1 ? c23 : c13), wy1 = p[1].y * (m == 1 ? c23 : c13);
        double wx2 = p[2].x * (m == 2 ? c23 : c13), wy2 = p[2].y * (m == 2 ? c23 : c13);
        double dxm = (wx0 + wx1) + wx2, dym = (wy0 + wy1) + wy2;
        double wp = (gx * dxm + gy * dym) + w;
        if (!is_wet(wp - p[m].z)) all_wet = false;
      }
      if (!all_wet) { gx = 0.; gy = 0.; }
      // TVD limiter on w (src/MUSCLObject.cpp:74-81)
      double tvd = 1.;
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        double vtmin = dmin(w, wn[kk]), vtmax = dmax(w, wn[kk]);
        double vek = w + (gx * (E[kk].x - c.x) + gy * (E[kk].y - c.y));
        if (!((vtmin <= vek) && (vek <= vtmax))) tvd = 0.;
      }
      gx = tvd * gx;
      gy = tvd * gy;
    }
  }
  W.mw[i] = w0;
  W.mgx[i] = gx;
  W.mgy[i] = gy;
  W.cls[i] = k;
}

// ---------------------------------------------------------------- pass 2 --
// ReconstructPartWetCell2 (src/MUSCLObject.cpp:114-191) for every partially
// wet cell (src/SpaceDisc.cpp:47-51).  The node maximum m_max_wp at the
// lowest vertex is max(z, AtPoint of each incident cell's pass-1 MUSCL).
__device__ inline double node_max_pass1(const DevMesh& M, const DevWork& W, int n) {
  V3 P = node(M, n);
  double mx = P.z;
  int b = M.n2c_off[n], e = M.n2c_off[n + 1];
  for (int q = b; q < e; ++q) {
    int c = M.n2c[q];
    V3 r[3];
    cell_nodes(M, c, r);
    V3 cc = centroid(r[0], r[1], r[2]);
    bool wet;
    double a = at_point_w(W.mw[c], W.mgx[c], W.mgy[c], cc, P, wet);
    mx = dmax(mx, a);
  }
  return mx;
}

template <bool HEAD_PW2>
__global__ void __launch_bounds__(256) k_recon2(DevMesh M, DevState S, DevWork W) {
  int64_t i = global_index();
  if (i >= M.T) return;
  if (W.cls[i] != 1) return;
  int ip[3] = {M.tp[i], M.tp[M.T + i], M.tp[2 * M.T + i]};
  V3 Q[3] = {node(M, ip[0]), node(M, ip[1]), node(M, ip[2])};
  V3 c = centroid(Q[0], Q[1], Q[2]);
  // sort by z with the reference's three conditional swaps
  {
    V3 t; int ti;
    if (Q[0].z > Q[1].z) { t = Q[0]; Q[0] = Q[1]; Q[1] = t; ti = ip[0]; ip[0] = ip[1]; ip[1] = ti; }
    if (Q[1].z > Q[2].z) { t = Q[1]; Q[1] = Q[2]; Q[2] = t; ti = ip[1]; ip[1] = ip[2]; ip[2] = ti; }
    if (Q[0].z > Q[1].z) { t = Q[0]; Q[0] = Q[1]; Q[1] = t; ti = ip[0]; ip[0] = ip[1]; ip[1] = ti; }
  }
  double w = S.w[i];
  double b23 = Q[0].z, b12 = Q[1].z, b13 = Q[2].z;
  double w0 = W.mw[i], gx = 0., gy = 0.;  // fallback: pass-1 (PartWet1) result
  bool fallback = (w > b13) || (b13 - b23 < kTol);
  if (!fallback) {
    double w23 = node_max_pass1(M, W, ip[0]);
    double h23 = w23 - b23;
    double ratio_b = (b12 - b23) / (b13 - b23);
    double hd1 = 1. / 3. * h23 * ratio_b;
    double hd2 = 1. / 3. * h23 * (2. * b13 - b12 - b23) / (b13 - b23);
    double hi = w - c.z;
    double ratio_h = hi / h23;
    V3 p0 = Q[0], p1, p2;
    if (!HEAD_PW2) p0.z = w23;
    if (hi <= hd1) {
      double k2 = sqrt(3. * ratio_h / ratio_b);
      p1 = {k2 * Q[1].x + (1. - k2) * Q[0].x, k2 * Q[1].y + (1. - k2) * Q[0].y,
            k2 * Q[1].z + (1. - k2) * Q[0].z};
      double k3 = sqrt(3. * ratio_h * ratio_b);
      p2 = {k3 * Q[2].x + (1. - k3) * Q[0].x, k3 * Q[2].y + (1. - k3) * Q[0].y,
            k3 * Q[2].z + (1. - k3) * Q[0].z};
    } else if (hi >= hd2) {
      double delta_w = 1.5 * (hi - hd2);
      p1 = Q[1];
      if (!HEAD_PW2) { p1.z += delta_w; p1.z += (1. - ratio_b) * h23; }
      p2 = Q[2];
      p2.z += delta_w;
    } else {
      double alpha = 3. * ratio_h;
      double beta = (b13 - b12) / (b13 - b23);
      double k1 = 1. - bisection(Cubic{0., (alpha - 3.) / beta, (1. + beta - alpha) / (beta * beta)},
                                 0., 1.);
      if (k1 < kTol) {
        fallback = true;
      } else {
        double k3 = 1. - beta * (1. - k1);
        double bp1 = k1 * b13 + (1. - k1) * b12;
        p1 = Q[1];
        if (!HEAD_PW2) p1.z = b12 + (k1 / k3) * beta * h23;
        p2 = {k1 * Q[2].x + (1. - k1) * Q[1].x, k1 * Q[2].y + (1. - k1) * Q[1].y,
              k1 * Q[2].z + (1. - k1) * Q[1].z};
        p2.z = bp1;
      }
    }
    if (!fallback) {
      gradient(p0, p1, p2, gx, gy);
      w0 = w23 + (gx * (c.x - Q[0].x) + gy * (c.y - Q[0].y));
    }
  }
  W.aw[i] = w0;
  W.agx[i] = gx;
  W.agy[i] = gy;
}

// ------------------------------------------------------------------ flux --
__device__ __forceinline__ void block_min_bits(double v, unsigned long long* target) {
  // wave-level min by DPP/shuffle, then one LDS slot per wave
  __shared__ double red[16];
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off, 64));
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    double m = red[0];
    for (int q = 1; q < nw; ++q) m = dmin(m, red[q]);
    // positive doubles order like their bit patterns
    if (m < 1.0) atomicMin(target, (unsigned long long)__double_as_longlong(m));
  }
}

// SpaceDisc::ComputeFluxes (src/SpaceDisc.cpp:54-74)
template <int FLUX, int WS>
__global__ void __launch_bounds__(256) k_flux(DevMesh M, DevState S, DevWork W, double cor,
                                              unsigned long long* rbits) {
  int64_t e = global_index();
  double lw = 1.0;
  if (e < M.E) {
    int a = M.ep0[e], b = M.ep1[e];
    V3 Pa = node(M, a), Pb = node(M, b);
    double le = edge_len(Pa, Pb);
    V3 E = midpoint(Pa, Pb);
    int f = M.et0[e], t = M.et1[e];
    V3 pf[3];
    cell_nodes(M, f, pf);
    V3 cf = centroid(pf[0], pf[1], pf[2]);
    double tx, ty;
    tangent(Pa, Pb, le, cf, tx, ty);
    double nx = ty, ny = -tx;
    double F[3];
    if (t < 0) {
      double h = S.w[f] - cf.z;
      elem_flux(nx, ny, h, 0., 0., F[0], F[1], F[2]);
    } else {
      V3 pt[3];
      cell_nodes(M, t, pt);
      V3 ct = centroid(pt[0], pt[1], pt[2]);
      double af = tri_area(pf[0], pf[1], pf[2]), at = tri_area(pt[0], pt[1], pt[2]);
      Slot L = slot_value(load_muscl(M, S, W, f), cf, E);
      Slot R = slot_value(load_muscl(M, S, W, t), ct, E);
      if (FLUX == 0) flux_hll<WS>(nx, ny, E.z, L, R, af, at, le, cor, lw, F);
      else flux_hllc<WS>(tx, ty, E.z, L, R, af, at, le, cor, lw, F);
    }
    W.f0[e] = F[0];
    W.f1[e] = F[1];
    W.f2[e] = F[2];
  }
  block_min_bits(lw, rbits);
}

// ----------------------------------------------------------- draining dt --
// TimeDisc::ComputeDrainingDt (src/TimeDisc.cpp:43-66) on the stage-start
// state.
__global__ void __launch_bounds__(256) k_drain(DevMesh M, DevState S, DevWork W) {
  int64_t i = global_index();
  if (i >= M.T) return;
  V3 p[3];
  cell_nodes(M, (int)i, p);
  V3 c = centroid(p[0], p[1], p[2]);
  double h = S.w[i] - c.z;
  double d;
  if (!is_wet(h)) {
    d = 0.;
  } else {
    uint32_t fl = M.cflag[i];
    double sum = 0.;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double fe = W.f0[M.te[k * M.T + i]];
      sum += dmax(0., sgn_pos(fl, k) ? fe : -fe);
    }
    d = sum > kTol ? tri_area(p[0], p[1], p[2]) * h / sum : INFINITY;
  }
  W.dtd[i] = d;
}

// -------------------------------------------------------------- update --
// TimeDisc::RHS (src/TimeDisc.cpp:3-41) + stage combination
// (src/Solvers.cpp:11-13,23-25,30-32,42-58) + ConsAssigner.
__device__ __forceinline__ void rhs_cell(const DevMesh& M, const DevState& S, const DevWork& W,
                                         int i, const V3 p[3], const V3& c, double dts,
                                         double cor, double res[3]) {
  uint32_t fl = M.cflag[i];
  Muscl mu = load_muscl(M, S, W, i);
  double i_area = 1. / tri_area(p[0], p[1], p[2]);
  double dti = W.dtd[i];
  res[0] = res[1] = res[2] = 0.;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int e = M.te[k * M.T + i];
    int nb = M.tt[k * M.T + i];
    double sg = sgn_pos(fl, k) ? 1. : -1.;
    double F0 = W.f0[e], F1 = W.f1[e], F2 = W.f2[e];
    double dtik = nb < 0 ? INFINITY : W.dtd[nb];
    double dtk = (sg * F0) > 0. ? dmin(dts, dti) : dmin(dts, dtik);
    const V3& Pa = p[ep0_local(fl, k)];
    const V3& Pb = p[ep1_local(fl, k)];
    double le = edge_len(Pa, Pb);
    double c_ek = i_area * le;
    V3 E = midpoint(Pa, Pb);
    Slot s = slot_value(mu, c, E);
    double h_ek = s.w - E.z;
    double src1 = mu.gx, src2 = mu.gy;
    src1 += cor * (-s.v);
    src2 += cor * s.u;
    double tx, ty;
    tangent(Pa, Pb, le, c, tx, ty);
    double nx = ty, ny = -tx;
    double cf = dtk * sg * c_ek;
    res[0] -= cf * F0;
    res[1] -= cf * F1;
    res[2] -= cf * F2;
    double cs = dts * (1. / 3.);
    res[1] -= cs * src1 * h_ek;
    res[2] -= cs * src2 * h_ek;
    double q = 0.5 * h_ek * h_ek;
    res[1] += dtk * (nx * c_ek * q);
    res[2] += dtk * (ny * c_ek * q);
  }
}

__device__ __forceinline__ double stage_dt(const StageArgs& A) {
  double dt = A.dt_fixed;
  if (!(dt > 0.)) dt = A.cfl * __longlong_as_double((long long)A.rbits[0]);
  return dt;
}

__global__ void __launch_bounds__(256) k_update(DevMesh M, DevState S, DevState U0, DevOut O,
                                                DevWork W, StageArgs A) {
  int64_t i = global_index();
  double dt = stage_dt(A);
  if (A.dts_log && blockIdx.x == 0 && threadIdx.x == 0) A.dts_log[0] = dt;
  if (i >= M.T) return;
  V3 p[3];
  cell_nodes(M, (int)i, p);
  V3 c = centroid(p[0], p[1], p[2]);
  double bc = c.z;
  double R[3];
  rhs_cell(M, S, W, (int)i, p, c, A.coef * dt, A.cor, R);
  double w = S.w[i], u = S.u[i], v = S.v[i];
  double h = w - bc;
  double Uc0 = h, Uc1 = h * u, Uc2 = h * v;
  double r0, r1, r2;
  if (A.a0 != 0.) {
    double w0 = U0.w[i], u0 = U0.u[i], v0 = U0.v[i];
    double h0 = w0 - bc;
    double Ua0 = h0, Ua1 = h0 * u0, Ua2 = h0 * v0;
    r0 = (A.a0 * Ua0 + A.a1 * Uc0) + R[0];
    r1 = (A.a0 * Ua1 + A.a1 * Uc1) + R[1];
    r2 = (A.a0 * Ua2 + A.a1 * Uc2) + R[2];
  } else {
    r0 = Uc0 + R[0];
    r1 = Uc1 + R[1];
    r2 = Uc2 + R[2];
  }
  double ow, ou, ov;
  cons_assign(bc, r0, r1, r2, ow, ou, ov);
  O.w[i] = ow;
  O.u[i] = ou;
  O.v[i] = ov;
}

__global__ void k_reset_r(unsigned long long* rbits) {
  *rbits = (unsigned long long)__double_as_longlong(1.0);
}

// ------------------------------------------------------------- getters --
// Edge and source slots as UpdateInterfaceValues leaves them
// (src/SpaceDisc.cpp:15-31), slot 2e + (from < to).
__global__ void __launch_bounds__(256) k_edge_field(DevMesh M, DevState S, DevWork W, double cor,
                                                    double* edg, double* src) {
  int64_t i = global_index();
  if (i >= M.T) return;
  V3 p[3];
  cell_nodes(M, (int)i, p);
  V3 c = centroid(p[0], p[1], p[2]);
  uint32_t fl = M.cflag[i];
  Muscl mu = load_muscl(M, S, W, (int)i);
  for (int k = 0; k < 3; ++k) {
    int e = M.te[k * M.T + i];
    int nb = M.tt[k * M.T + i];
    int64_t s = 2 * (int64_t)e + (((int)i < nb) ? 1 : 0);
    V3 E = midpoint(p[ep0_local(fl, k)], p[ep1_local(fl, k)]);
    Slot sv = slot_value(mu, c, E);
    edg[3 * s] = sv.w;
    edg[3 * s + 1] = sv.u;
    edg[3 * s + 2] = sv.v;
    double s1 = mu.gx, s2 = mu.gy;
    s1 += cor * (-sv.v);
    s2 += cor * sv.u;
    src[3 * s] = 0.;
    src[3 * s + 1] = s1;
    src[3 * s + 2] = s2;
  }
}

// m_max_wp after ComputeInterfaceValues: max of z and AtPoint of pass-1 and
// (partially wet cells) pass-2 reconstructions at the node.
__global__ void __launch_bounds__(256) k_max_wp(DevMesh M, DevWork W, double* out) {
  int64_t n = global_index();
  if (n >= M.V) return;
  V3 P = node(M, (int)n);
  double mx = P.z;
  for (int q = M.n2c_off[n]; q < M.n2c_off[n + 1]; ++q) {
    int c = M.n2c[q];
    V3 r[3];
    cell_nodes(M, c, r);
    V3 cc = centroid(r[0], r[1], r[2]);
    bool wet;
    mx = dmax(mx, at_point_w(W.mw[c], W.mgx[c], W.mgy[c], cc, P, wet));
    if (W.cls[c] == 1) mx = dmax(mx, at_point_w(W.aw[c], W.agx[c], W.agy[c], cc, P, wet));
  }
  out[n] = mx;
}

__global__ void k_rhs_one(DevMesh M, DevState S, DevWork W, int i, double dt, double cor,
                          double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  V3 p[3];
  cell_nodes(M, i, p);
  V3 c = centroid(p[0], p[1], p[2]);
  rhs_cell(M, S, W, i, p, c, dt, cor, out);
}

// -------------------------------------------------------------- launch --
static inline dim3 grid_for(int64_t n) {
  int64_t nb = (n + 255) / 256;
  nb = ((nb + 7) / 8) * 8;
  if (nb < 8) nb = 8;
  return dim3((unsigned)nb);
}

hipError_t launch_recon1(const DevMesh& M, const DevState& S, const DevWork& W, int opts,
                         hipStream_t st) {
  if (opts & 1) hipLaunchKernelGGL((k_recon1<true>), grid_for(M.T), dim3(256), 0, st, M, S, W);
  else hipLaunchKernelGGL((k_recon1<false>), grid_for(M.T), dim3(256), 0, st, M, S, W);
  return hipGetLastError();
}
hipError_t launch_recon2(const DevMesh& M, const DevState& S, const DevWork& W, int opts,
                         hipStream_t st) {
  if (opts & 2) hipLaunchKernelGGL((k_recon2<true>), grid_for(M.T), dim3(256), 0, st, M, S, W);
  else hipLaunchKernelGGL((k_recon2<false>), grid_for(M.T), dim3(256), 0, st, M, S, W);
  return hipGetLastError();
}
hipError_t launch_flux(const DevMesh& M, const DevState& S, const DevWork& W, int flux, int ws,
                       double cor, unsigned long long* rbits, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_r, dim3(1), dim3(1), 0, st, rbits);
  dim3 g = grid_for(M.E), b(256);
#define SWE_FLUX_CASE(F, Wv)                                                           \
  if (flux == F && ws == Wv) {                                                         \
    hipLaunchKernelGGL((k_flux<F, Wv>), g, b, 0, st, M, S, W, cor, rbits);             \
    return hipGetLastError();                                                          \
  }
  SWE_FLUX_CASE(0, 0) SWE_FLUX_CASE(0, 1) SWE_FLUX_CASE(0, 2)
  SWE_FLUX_CASE(1, 0) SWE_FLUX_CASE(1, 1) SWE_FLUX_CASE(1, 2)
#undef SWE_FLUX_CASE
  return hipErrorInvalidValue;
}
hipError_t launch_drain(const DevMesh& M, const DevState& S, const DevWork& W, hipStream_t st) {
  hipLaunchKernelGGL(k_drain, grid_for(M.T), dim3(256), 0, st, M, S, W);
  return hipGetLastError();
}
hipError_t launch_update(const DevMesh& M, const DevState& S, const DevState& U0, const DevOut& O,
                         const DevWork& W, const StageArgs& A, hipStream_t st) {
  hipLaunchKernelGGL(k_update, grid_for(M.T), dim3(256), 0, st, M, S, U0, O, W, A);
  return hipGetLastError();
}
hipError_t launch_edge_field(const DevMesh& M, const DevState& S, const DevWork& W, double cor,
                             double* edg, double* src, hipStream_t st) {
  hipLaunchKernelGGL(k_edge_field, grid_for(M.T), dim3(256), 0, st, M, S, W, cor, edg, src);
  return hipGetLastError();
}
hipError_t launch_max_wp(const DevMesh& M, const DevWork& W, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_max_wp, grid_for(M.V), dim3(256), 0, st, M, W, out);
  return hipGetLastError();
}
hipError_t launch_rhs_one(const DevMesh& M, const DevState& S, const DevWork& W, int i, double dt,
                          double cor, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_rhs_one, dim3(1), dim3(64), 0, st, M, S, W, i, dt, cor, out);
  return hipGetLastError();
}

}  // namespace swe
