// Helpers shared by the step kernels (swe_kernels.hip) and the adjoint
// kernels (swe_adjoint.hip): launch geometry, per-cell flag decoding, node
// access, the final per-cell MUSCL and the edge slot it induces.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swe_device.h"
#include "swe_launch.h"

namespace swe {

#define CELL_BND (1u << 3)
// boundary tags (include/Includes.h:21): walls (src/SpaceDisc.cpp:66-69) and
// free flow, whose outside state is the inside edge state (transmissive)
constexpr int kSolidWall = -1;
// PERIODIC: the outside state is the partner edge's inside slot (M.pe)
constexpr int kPeriodic = -3;
constexpr int kOptDrainLength = 32;  // SWE_OPT_DRAIN_LENGTH

__device__ __forceinline__ int logical_block() {
  // Workgroups are dispatched round-robin over the 8 XCDs; hand each XCD a
  // contiguous range of the Morton-ordered cells so neighbour gathers hit
  // that XCD's L2.  gridDim.x is a multiple of 8.
  int per = gridDim.x >> 3;
  int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}
// The same ranges, each walked from its far end.  A kernel launched after
// one that walked them forward starts on the lines the other touched last,
// which the Infinity Cache still holds (k_flux after the reconstruction,
// the evaluating tile kernel after the second reconstruction).
__device__ __forceinline__ int logical_block_reversed() {
  int per = gridDim.x >> 3;
  int b = blockIdx.x;
  return (b & 7) * per + (per - 1 - (b >> 3));
}
constexpr int kBlock = 256;  // every launch uses 256-thread workgroups
__device__ __forceinline__ int64_t global_index() {
  return (int64_t)logical_block() * kBlock + threadIdx.x;
}

__device__ __forceinline__ V3 node(const DevMesh& M, int n) {
  const NodeRec r = M.pn[n];
  return {r.x, r.y, r.z};
}
__device__ __forceinline__ int ep0_local(uint32_t fl, int k) { return (fl >> (4 + 4 * k)) & 3; }
__device__ __forceinline__ int ep1_local(uint32_t fl, int k) { return (fl >> (6 + 4 * k)) & 3; }
__device__ __forceinline__ bool sgn_pos(uint32_t fl, int k) { return (fl >> k) & 1u; }
__device__ __forceinline__ V3 pick(const V3& a, const V3& b, const V3& c, int i) {
  return i == 0 ? a : (i == 1 ? b : c);
}
// midpoint of edge k from the cell's own vertices
__device__ __forceinline__ V3 edge_mid(const V3& p0, const V3& p1, const V3& p2, uint32_t fl, int k) {
  return midpoint(pick(p0, p1, p2, ep0_local(fl, k)), pick(p0, p1, p2, ep1_local(fl, k)));
}

// Final MUSCL of cell c (pass 2 has overwritten the partially wet cells in
// place); every load is issued unconditionally so they overlap.
struct Muscl {
  double w0, gx, gy, u0, v0;
};
__device__ __forceinline__ Muscl load_muscl(const DevState& S, const DevWork& W, int c) {
  const int k = W.cls[c];
  const double w0 = W.mw[c], gx = W.mgx[c], gy = W.mgy[c];
  const double u = S.u[S.at(c)], v = S.v[S.at(c)];
  Muscl m{w0, gx, gy, u, v};
  if (k == 0) {
    m.u0 = 0.; m.v0 = 0.;
  }
  return m;
}

// AtPoint + PrimAssigner: the edge slot written by UpdateInterfaceValues
// (src/SpaceDisc.cpp:24-25); (cx, cy) is the cell centroid.
__device__ __forceinline__ Slot slot_value(const Muscl& m, double cx, double cy, const V3& E) {
  double dx = E.x - cx, dy = E.y - cy;
  double aw = m.w0 + (m.gx * dx + m.gy * dy);
  bool wet = (aw - E.z) >= 0;
  Slot s;
  if (!wet) aw = E.z;
  prim_assign(E.z, aw, wet ? m.u0 : 0., wet ? m.v0 : 0., s.w, s.u, s.v);
  return s;
}

// TimeDisc::RHS (src/TimeDisc.cpp:3-41) of one cell from its inputs: its
// MUSCL, vertices and draining time step (RhsCell), and per local edge k
// the edge's flux and the neighbour's draining time step, which the
// accessor edge(k, F0, F1, F2, dtik) supplies (from registers in the
// update kernel, from LDS in the fused kernel: the same arithmetic).
struct RhsCell {
  uint32_t fl;
  Muscl mu;
  double wi;  // cell average level
  V3 p[3];  // the cell's vertices
  double dti;
};
struct RhsIn : RhsCell {
  double F0[3], F1[3], F2[3], dtik[3];
};
// eb (SWE_OPT_EXACT_BALANCE): each edge adds the flux minus the cell's own
// hydrostatic edge pressure, sg F - n h_ek^2 / 2, which is zero to the bit
// where the edge states on both sides are equal and at rest
template <class EdgeIn>
__device__ __forceinline__ void rhs_compute_by(const RhsCell& q, EdgeIn edge, double dts, double cor,
                                               double fric, bool eb, double& r0, double& r1, double& r2) {
  r0 = 0.; r1 = 0.; r2 = 0.;
  // Domain::T / Area (src/Bathymetry.cpp:24-27,84-90)
  const V3 c = centroid(q.p[0], q.p[1], q.p[2]);
  const double i_area = 1. / tri_area(q.p[0], q.p[1], q.p[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bool own = sgn_pos(q.fl, k);
    const double sg = own ? 1. : -1.;
    double F0, F1, F2, dtik;
    edge(k, F0, F1, F2, dtik);
    const double dtk = (sg * F0) > 0. ? dmin(dts, q.dti) : dmin(dts, dtik);
    // Domain::L / E / Tang of edge k (src/Bathymetry.cpp:29-76); the tangent
    // is taken outward of this cell, as Norm(e, i) is
    const V3 Pa = pick(q.p[0], q.p[1], q.p[2], ep0_local(q.fl, k));
    const V3 Pb = pick(q.p[0], q.p[1], q.p[2], ep1_local(q.fl, k));
    const double le = edge_len(Pa, Pb);
    const double c_ek = i_area * le;
    const V3 E = midpoint(Pa, Pb);
    const Slot s = slot_value(q.mu, c.x, c.y, E);
    const double h_ek = s.w - E.z;
    double src1 = q.mu.gx, src2 = q.mu.gy;
    src1 += cor * (-s.v);
    src2 += cor * s.u;
    double tx, ty;
    tangent(Pa, Pb, le, c, tx, ty);
    const double nx = ty, ny = -tx;
    if (eb) {
      const double q2 = 0.5 * h_ek * h_ek;
      const double g0 = sg * F0, g1 = sg * F1 - nx * q2, g2 = sg * F2 - ny * q2;
      r0 -= dtk * (c_ek * g0);
      r1 -= dtk * (c_ek * g1);
      r2 -= dtk * (c_ek * g2);
      const double cs = dts * (1. / 3.);
      r1 -= cs * src1 * h_ek;
      r2 -= cs * src2 * h_ek;
      continue;
    }
    const double cf = dtk * sg * c_ek;
    r0 -= cf * F0;
    r1 -= cf * F1;
    r2 -= cf * F2;
    const double cs = dts * (1. / 3.);
    r1 -= cs * src1 * h_ek;
    r2 -= cs * src2 * h_ek;
    const double q2 = 0.5 * h_ek * h_ek;
    r1 += dtk * (nx * c_ek * q2);
    r2 += dtk * (ny * c_ek * q2);
  }
  // linear friction -tau (hu, hv) of the cell average (opt-in; the
  // reference stores tau unused, include/SpaceDisc.h:47)
  if (fric != 0.) {
    const double hh = q.wi - c.z;
    r1 -= dts * (fric * (hh * q.mu.u0));
    r2 -= dts * (fric * (hh * q.mu.v0));
  }
}
// stage combination a0 U0 + a1 Uc + R (src/Solvers.cpp:23-25,42-58); cvx
// (SWE_OPT_EXACT_BALANCE with a0 + a1 = 1) writes it Uc + a0 (U0 - Uc) + R,
// which is Uc + R to the bit where U0 = Uc
__device__ __forceinline__ double stage_combine(bool cvx, double a0, double a1, double Ua, double Uc, double R) {
  if (cvx) return (Uc + a0 * (Ua - Uc)) + R;
  return (a0 * Ua + a1 * Uc) + R;
}
__device__ __forceinline__ void rhs_compute(const RhsIn& q, double dts, double cor, double fric, bool eb,
                                            double& r0, double& r1, double& r2) {
  rhs_compute_by(
      q,
      [&](int k, double& F0, double& F1, double& F2, double& dtik) {
        F0 = q.F0[k]; F1 = q.F1[k]; F2 = q.F2[k]; dtik = q.dtik[k];
      },
      dts, cor, fric, eb, r0, r1, r2);
}
// grid of 256-thread workgroups covering n items, a multiple of 8 (one
// contiguous range per XCD, see logical_block)
// a store of a kernel output; NT (a template parameter: a runtime branch
// between the two forms is merged into a plain store) chosen at launch from
// DevMesh::nt
template <bool NT>
__device__ __forceinline__ void out_store(double* p, double v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

inline dim3 grid_for_bs(int64_t n, int64_t bs) {
  int64_t nb = (n + bs - 1) / bs;
  nb = ((nb + 7) / 8) * 8;
  if (nb < 8) nb = 8;
  return dim3((unsigned)nb);
}
inline dim3 grid_for(int64_t n) { return grid_for_bs(n, kBlock); }

}  // namespace swe
