// gfx950 kernels of the explicit finite-volume step.
//
// Per stage (Solvers::Euler/SSPRK2/SSPRK3 stage, src/Solvers.cpp:5-59):
//   k_recon1  per cell  — classification + pass-1 MUSCL (dry / partially wet
//                         level / full-wet limited plane), stored as
//                         (w0, gx, gy) per cell: the reference's edge and
//                         source slots are functions of it and are never
//                         materialised on the hot path.  Partially wet cells
//                         are appended to a work list (one atomic per wave).
//   k_recon2  per listed cell — pass 2 for partially wet cells (Liu et al.
//                         plane) with the node maximum gathered over the
//                         node->cell map (no atomics, order-independent max).
//   k_flux    per edge  — HLL/HLLC numerical flux from the two cells' MUSCL
//                         at the edge midpoint; wave-level + block min of the
//                         CFL length-to-wavespeed, one 64-bit atomic min per
//                         block on the positive-double bit pattern.
//   k_drain   per cell  — draining time step from the stage-start state.
//   k_update  per cell  — fixed-order gather of the three edge fluxes, the
//                         well-balanced source, SSP-RK stage combination and
//                         ConsAssigner write-back.
// Cell geometry that neighbours gather (centroid, bathymetry range, area) and
// the flux's CFL length are computed once by the setup kernels; the edge
// geometry (length, midpoint, tangent) and the updated cell's own centroid and
// area are recomputed from the node coordinates with the same device
// functions (bit-identical), which costs less than streaming them from HBM.
// Every reduction is either order-independent (max, min) or a fixed-order
// per-cell gather, so results are deterministic run to run.  Local arrays
// are only indexed by compile-time constants (unrolled loops) or through
// register selects, so nothing is placed in scratch or LDS.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "swe_common.h"

namespace swe {

// -------------------------------------------------------------- setup --
// Static geometry with the formulas of Domain (src/Bathymetry.cpp:24-90).
// the interleaved node records, before the cell and edge setup reads them
__global__ void __launch_bounds__(256) k_setup_nodes(DevMesh M, DevGeomOut G) {
  int64_t n = global_index();
  if (n >= M.V) return;
  G.pn[n] = NodeRec{M.px[n], M.py[n], M.pz[n], 0.};
}

__global__ void __launch_bounds__(256) k_setup_cells(DevMesh M, DevGeomOut G) {
  int64_t i = global_index();
  if (i >= M.T) return;
  V3 p0 = node(M, M.tp[i]), p1 = node(M, M.tp[M.T + i]), p2 = node(M, M.tp[2 * M.T + i]);
  V3 c = centroid(p0, p1, p2);
  G.cx[i] = c.x;
  G.cy[i] = c.y;
  G.bc[i] = c.z;
  G.area[i] = tri_area(p0, p1, p2);
  CellRec r;
  r.cx = c.x; r.cy = c.y; r.bc = c.z;
  r.zmin = dmin(dmin(p0.z, p1.z), p2.z);
  r.zmax = dmax(dmax(p0.z, p1.z), p2.z);
  r.fl = M.cflag[i];
  r.pad = 0;
  G.crec[i] = r;
  G.nrec[i] = NRec{r.cx, r.cy, (r.fl & CELL_BND) ? INFINITY : r.zmax, r.bc};
  G.cxy[i] = CXY{c.x, c.y};
  G.cv[i] = CVtx{M.tp[i], M.tp[M.T + i], M.tp[2 * M.T + i], M.cflag[i]};
}

// min over the edge's two cells of 2 area / length (the flux's CFL length)
__global__ void __launch_bounds__(256) k_setup_edges(DevMesh M, DevGeomOut G) {
  int64_t e = global_index();
  if (e >= M.E) return;
  const int f = M.et0[e], t = M.et1[e];
  double dm = 0.;
  if (t != kSolidWall) {  // interior, or free flow (the inside cell twice)
    const double le = edge_len(node(M, M.ep0[e]), node(M, M.ep1[e]));
    const double dl = 2. * G.area[f] / le;
    double dr = t < 0 ? dl : 2. * G.area[t] / le;
    if (t == kPeriodic) dr = 2. * G.area[M.et0[M.pe[e]]] / le;
    dm = dmin(dl, dr);
  }
  G.edmin[e] = dm;
}

// ReconstructFullWetCell (src/MUSCLObject.cpp:38-84): the limited plane of
// cell (vertices p0..p2, flags fl, level w, centroid cx, cy) from its
// neighbours' levels wn, records rn and ids jn (a partially wet neighbour's
// full record is read from M.crec).
template <bool HEAD_GRAD>
__device__ __forceinline__ void fullwet_plane(const DevMesh& M, const V3& p0, const V3& p1, const V3& p2,
                                              uint32_t fl, double w, double cx, double cy, const double wn[3],
                                              const NRec rn[3], const int jn[3], double& gx, double& gy) {
  gx = 0.;
  gy = 0.;
  V3 pts[3];
  V3 E[3];
  bool any_dry = false;
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    E[kk] = edge_mid(p0, p1, p2, fl, kk);
    const double wj = wn[kk], bcj = rn[kk].bc;
    if (rn[kk].zmx < wj) {  // interior full-wet neighbour (zmx = +inf on boundary cells)
      pts[kk] = {rn[kk].cx, rn[kk].cy, HEAD_GRAD ? bcj : wj};
    } else if (!is_wet(wj - bcj)) {
      any_dry = true;
      pts[kk] = {0., 0., 0.};
    } else {
      const CellRec rj = M.crec[jn[kk]];  // partially wet neighbour: the full record
      double w0j = partwet1_w(rj.zmin, rj.zmin, rj.zmax, bcj, wj);
      double dx = E[kk].x - rn[kk].cx, dy = E[kk].y - rn[kk].cy;
      double a = w0j + (0. * dx + 0. * dy);
      if (!((a - E[kk].z) >= 0)) a = E[kk].z;
      double val = 0.5 * (w + a);
      pts[kk] = {E[kk].x, E[kk].y, HEAD_GRAD ? E[kk].z : val};
    }
  }
  if (any_dry) return;
  gradient(pts[0], pts[1], pts[2], gx, gy);
  // vertex positivity (src/MUSCLObject.cpp:66-72)
  const double c23 = 1. - 1. / 3., c13 = -1. / 3.;
  bool all_wet = true;
  {
    double dxm = (p0.x * c23 + p1.x * c13) + p2.x * c13, dym = (p0.y * c23 + p1.y * c13) + p2.y * c13;
    if (!is_wet(((gx * dxm + gy * dym) + w) - p0.z)) all_wet = false;
  }
  {
    double dxm = (p0.x * c13 + p1.x * c23) + p2.x * c13, dym = (p0.y * c13 + p1.y * c23) + p2.y * c13;
    if (!is_wet(((gx * dxm + gy * dym) + w) - p1.z)) all_wet = false;
  }
  {
    double dxm = (p0.x * c13 + p1.x * c13) + p2.x * c23, dym = (p0.y * c13 + p1.y * c13) + p2.y * c23;
    if (!is_wet(((gx * dxm + gy * dym) + w) - p2.z)) all_wet = false;
  }
  if (!all_wet) { gx = 0.; gy = 0.; }
  // TVD limiter on w (src/MUSCLObject.cpp:74-81)
  double tvd = 1.;
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    double vtmin = dmin(w, wn[kk]), vtmax = dmax(w, wn[kk]);
    double vek = w + (gx * (E[kk].x - cx) + gy * (E[kk].y - cy));
    if (!((vtmin <= vek) && (vek <= vtmax))) tvd = 0.;
  }
  gx = tvd * gx;
  gy = tvd * gy;
}

// ---------------------------------------------------------------- pass 1 --
// SpaceDisc::ComputeInterfaceValues pass 1 (src/SpaceDisc.cpp:37-45) with
// ReconstructDryCell / ReconstructPartWetCell1 / ReconstructFullWetCell
// (src/MUSCLObject.cpp:31-112).
template <bool HEAD_GRAD, bool NT>
__global__ void __launch_bounds__(256) k_recon1(DevMesh M, DevState S, DevWork W) {
  // The workgroup's 256 Morton-consecutive cells stage their level and the
  // part of their record a neighbour reads (NRec) in LDS; most neighbours
  // lie in the same workgroup and are read from there, the others from
  // global memory.
  __shared__ double s_w[kBlock];
  __shared__ NRec s_r[kBlock];
  const int64_t blk0 = M.c0 + (int64_t)logical_block() * kBlock;
  const int64_t i = blk0 + threadIdx.x;
  const bool live = i < M.c1;
  const int is = live ? (int)i : (int)M.c0;  // threads past the end read a valid cell
  // every load that does not need the LDS window is issued before the
  // barrier: the cell's own record (which fills its LDS entry), its flags,
  // vertices and neighbour ids, and the neighbours outside the workgroup
  const NRec ri = M.nrec[is];
  const double w = S.w[S.at(is)];
  const CVtx cvi = M.cv[is];
  const uint32_t fl = cvi.fl;
  int jn[3];
  bool inb[3];
  double wn[3];
  NRec rn[3];
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    int j = M.tt[kk * M.T + is];
    j = j < 0 ? is : j;  // boundary cells never take the full-wet branch
    jn[kk] = j;
    const int64_t lj = j - blk0;
    inb[kk] = lj >= 0 && lj < kBlock && j < M.c1;
    if (!inb[kk]) {  // the owner's workgroup reads the same record: an L2 hit mostly
      wn[kk] = S.w[S.at(j)];
      rn[kk] = M.nrec[j];
    }
  }
  const V3 p0 = node(M, cvi.v0), p1 = node(M, cvi.v1), p2 = node(M, cvi.v2);
  if (live) {
    s_w[threadIdx.x] = w;
    s_r[threadIdx.x] = ri;
  }
  __syncthreads();
  if (!live) return;
  if (M.cinst && !M.bactive[M.cinst[i]]) return;  // finished ensemble member
#pragma unroll
  for (int kk = 0; kk < 3; ++kk)
    if (inb[kk]) {
      const int64_t lj = jn[kk] - blk0;
      wn[kk] = s_w[lj];
      rn[kk] = s_r[lj];
    }
  const double bc = ri.bc;
  const double cx = ri.cx, cy = ri.cy;
  double w0 = w, gx = 0., gy = 0.;
  double zmx = 0.;  // the highest vertex, read where pass 1 or 2 needs it
  int8_t k;
  if (!is_wet(w - bc)) {
    k = 0;
    w0 = bc;
    gradient(p0, p1, p2, gx, gy);
  } else if (!(ri.zmx < w)) {  // a boundary cell (zmx = +inf) or a vertex at or above w
    k = 1;
    // b13 = zmax, b23 = zmin enter partwet1_w only through max / min
    const CellRec rf = M.crec[i];
    zmx = rf.zmax;
    w0 = partwet1_w(rf.zmin, rf.zmin, zmx, bc, w);
    W.wrec1[i] = w0;
  } else {
    k = 2;
    fullwet_plane<HEAD_GRAD>(M, p0, p1, p2, fl, w, cx, cy, wn, rn, jn, gx, gy);
  }
  // outputs read only by later kernels: out_store, nontemporal on large
  // meshes, keeps L2 for the neighbour gathers (-1.5 % per step at 16M cells)
  out_store<NT>(&W.mw[i], w0);
  out_store<NT>(&W.mgx[i], gx);
  out_store<NT>(&W.mgy[i], gy);
  W.cls[i] = k;
  // append partially wet cells to pass 2's work list, one atomic per wave
  // (the order of the list is irrelevant: pass 2 is independent per cell).
  // A level above every vertex (wet boundary cells) is pass 2's fallback to
  // the pass-1 level already written: such cells are left out.
  const bool pw = k == 1 && !(w > zmx);
  const unsigned long long m = __ballot(pw);
  if (m) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(W.pw_count, (unsigned int)__popcll(m));
    base = __shfl(base, leader, 64);
    if (pw) {
      const unsigned int below = __builtin_amdgcn_mbcnt_hi(
          (unsigned int)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)m, 0u));
      W.pw_list[base + below] = (int)i;
    }
  }
}

// ---------------------------------------------------------------- pass 2 --
// ReconstructPartWetCell2 (src/MUSCLObject.cpp:114-191) for every partially
// wet cell (src/SpaceDisc.cpp:47-51).  The node maximum m_max_wp at the
// lowest vertex is max(z, AtPoint of each incident cell's pass-1 MUSCL).
__device__ inline double node_max_pass1(const DevMesh& M, const DevWork& W, int n, const V3& P) {
  double mx = P.z;
  int b = M.n2c_off[n], e = M.n2c_off[n + 1];
  for (int q = b; q < e; ++q) {
    const int c = M.n2c[q];
    const int k = W.cls[c];
    // pass-1 reconstruction of c: partially wet cells hold their pass-1
    // level in wrec1 (zero gradient) while pass 2 rewrites mw in place
    double w0 = W.mw[c], gx = W.mgx[c], gy = W.mgy[c];
    if (k == 1) { w0 = W.wrec1[c]; gx = 0.; gy = 0.; }
    double dx = P.x - M.cx[c], dy = P.y - M.cy[c];
    double a = w0 + (gx * dx + gy * dy);
    if (!((a - P.z) >= 0)) a = P.z;
    mx = dmax(mx, a);
  }
  return mx;
}

// the final MUSCL (w0, gx, gy) of a partially wet cell i; w1: its pass-1
// (PartWet1) level, the result where pass 2 falls back to it
template <bool HEAD_PW2>
__device__ inline void partwet2_muscl(const DevMesh& M, const DevState& S, const DevWork& W, int i, double w1,
                                      double& w0_out, double& gx_out, double& gy_out) {
  int i0 = M.tp[i], i1 = M.tp[M.T + i], i2 = M.tp[2 * M.T + i];
  V3 Q0 = node(M, i0), Q1 = node(M, i1), Q2 = node(M, i2);
  const double cx = M.cx[i], cy = M.cy[i], bc = M.bc[i];
  // sort by z with the reference's three conditional swaps
  {
    V3 t; int ti;
    if (Q0.z > Q1.z) { t = Q0; Q0 = Q1; Q1 = t; ti = i0; i0 = i1; i1 = ti; }
    if (Q1.z > Q2.z) { t = Q1; Q1 = Q2; Q2 = t; ti = i1; i1 = i2; i2 = ti; }
    if (Q0.z > Q1.z) { t = Q0; Q0 = Q1; Q1 = t; ti = i0; i0 = i1; i1 = ti; }
  }
  (void)i1;
  (void)i2;
  double w = S.w[S.at(i)];
  double b23 = Q0.z, b12 = Q1.z, b13 = Q2.z;
  double w0 = w1, gx = 0., gy = 0.;  // fallback: pass-1 (PartWet1) result
  bool fallback = (w > b13) || (b13 - b23 < kTol);
  if (!fallback) {
    double w23 = node_max_pass1(M, W, i0, Q0);
    double h23 = w23 - b23;
    double ratio_b = (b12 - b23) / (b13 - b23);
    double hd1 = 1. / 3. * h23 * ratio_b;
    double hd2 = 1. / 3. * h23 * (2. * b13 - b12 - b23) / (b13 - b23);
    double hi = w - bc;
    double ratio_h = hi / h23;
    V3 p0 = Q0, p1, p2;
    if (!HEAD_PW2) p0.z = w23;
    if (hi <= hd1) {
      double k2 = sqrt(3. * ratio_h / ratio_b);
      p1 = {k2 * Q1.x + (1. - k2) * Q0.x, k2 * Q1.y + (1. - k2) * Q0.y, k2 * Q1.z + (1. - k2) * Q0.z};
      double k3 = sqrt(3. * ratio_h * ratio_b);
      p2 = {k3 * Q2.x + (1. - k3) * Q0.x, k3 * Q2.y + (1. - k3) * Q0.y, k3 * Q2.z + (1. - k3) * Q0.z};
    } else if (hi >= hd2) {
      double delta_w = 1.5 * (hi - hd2);
      p1 = Q1;
      if (!HEAD_PW2) { p1.z += delta_w; p1.z += (1. - ratio_b) * h23; }
      p2 = Q2;
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
        p1 = Q1;
        if (!HEAD_PW2) p1.z = b12 + (k1 / k3) * beta * h23;
        p2 = {k1 * Q2.x + (1. - k1) * Q1.x, k1 * Q2.y + (1. - k1) * Q1.y, k1 * Q2.z + (1. - k1) * Q1.z};
        p2.z = bp1;
      }
    }
    if (!fallback) {
      gradient(p0, p1, p2, gx, gy);
      w0 = w23 + (gx * (cx - Q0.x) + gy * (cy - Q0.y));
    }
  }
  w0_out = w0;
  gx_out = gx;
  gy_out = gy;
}

template <bool HEAD_PW2>
__device__ inline void recon2_cell(const DevMesh& M, const DevState& S, const DevWork& W, int i) {
  double w0, gx, gy;
  partwet2_muscl<HEAD_PW2>(M, S, W, i, W.wrec1[i], w0, gx, gy);
  // in place: readers of pass-1 values take partially wet cells from wrec1
  W.mw[i] = w0;
  W.mgx[i] = gx;
  W.mgy[i] = gy;
}

// over the list of partially wet cells written by k_recon1
template <bool HEAD_PW2>
__global__ void __launch_bounds__(256) k_recon2(DevMesh M, DevState S, DevWork W) {
  const unsigned int n = *W.pw_count;
  for (unsigned int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock)
    recon2_cell<HEAD_PW2>(M, S, W, W.pw_list[q]);
}
__global__ void k_zero_count(unsigned int* c) { *c = 0u; }

// ------------------------------------------------------------------ flux --
__device__ __forceinline__ void block_min_bits(double v, unsigned long long* target) {
  // wave-level min by cross-lane shuffles, then one LDS slot per wave
  __shared__ double red[16];
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off, 64));
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
    double m = red[0];
    for (int q = 1; q < nw; ++q) m = dmin(m, red[q]);
    // positive doubles order like their bit patterns
    if (m < 1.0) atomicMin(target, (unsigned long long)__double_as_longlong(m));
  }
}

// SpaceDisc::ComputeFluxes (src/SpaceDisc.cpp:54-74).  Each thread handles
// NI edges BS apart (BS threads per workgroup); all their loads are issued
// before any flux is computed, so a wave keeps NI times as many gathers in
// flight.  NI and BS are launch-time choices (swe_autotune); the results do
// not depend on them.
// CFL false: no length-to-wavespeed minimum (a fixed-dt pass that never
// reads it); the fluxes are the same.
struct FluxIn {
  V3 Pa, Pb;  // P(ep0), P(ep1)
  V3 Eq;      // periodic: the partner edge's midpoint (the right state's point)
  double edmin, hf, cxf, cyf, cxt, cyt;
  Muscl mf, mt;
  int64_t e;
  int t, inst;
  bool on, neg;  // neg: the second edge of a periodic pair takes -F of the first
};
template <int FLUX, int WS, int NI, bool CFL, int BS, bool NT>
__global__ void __launch_bounds__(BS) k_flux(DevMesh M, DevState S, DevWork W, double cor,
                                              unsigned long long* rbits) {
  // edges walked backwards: the reconstruction before walks forwards
  const int64_t base = M.e0 + (int64_t)logical_block_reversed() * (BS * NI) + threadIdx.x;
  FluxIn in[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int64_t e = base + it * BS;
    const int64_t es = e < M.e1 ? e : M.e0;  // out-of-range items read a valid edge and are dropped
    FluxIn& q = in[it];
    q.e = e;
    int f = M.et0[es];
    const int t = M.et1[es];
    int tr = t < 0 ? f : t;  // boundary edges read their own cell twice: R = L
    int64_t eg = es;         // the edge whose geometry and orientation the flux takes
    q.neg = false;
    if (t == kPeriodic) {  // the pair's first edge, its cell left, the partner's right
      const int qe = M.pe[es];
      const int p = es < qe ? (int)es : qe, s2 = es < qe ? qe : (int)es;
      eg = p;
      f = M.et0[p];
      tr = M.et0[s2];
      q.neg = es != p;
      q.Eq = midpoint(node(M, M.ep0[s2]), node(M, M.ep1[s2]));
    }
    q.Pa = node(M, M.ep0[eg]);
    q.Pb = node(M, M.ep1[eg]);
    q.edmin = CFL ? M.edmin[eg] : 1.;
    q.t = t;
    q.mf = load_muscl(S, W, f);
    q.mt = load_muscl(S, W, tr);
    const CXY cf = M.cxy[f], ct = M.cxy[tr];
    q.cxf = cf.x;
    q.cyf = cf.y;
    q.cxt = ct.x;
    q.cyt = ct.y;
    // the cell average enters only the wall flux: walls alone load it
    // (with SWE_OPT_EXACT_BALANCE the wall takes the inside edge depth)
    q.hf = (t == kSolidWall && !M.eb) ? S.w[S.at(f)] - M.bc[f] : 0.;
    q.inst = M.cinst ? M.cinst[f] : 0;
    q.on = e < M.e1 && (!M.cinst || M.bactive[q.inst]);
  }
  double lw = 1.0;
  double lwi_item[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const FluxIn& q = in[it];
    lwi_item[it] = 1.0;
    if (!q.on) continue;
    // Domain::L / E / Tang (src/Bathymetry.cpp:29-76), outward of et0
    const double le = edge_len(q.Pa, q.Pb);
    const V3 E = midpoint(q.Pa, q.Pb);
    double tx, ty;
    tangent(q.Pa, q.Pb, le, V3{q.cxf, q.cyf, 0.}, tx, ty);
    const double nx = ty, ny = -tx;
    double F[3];
    double lwi = 1.0;
    if (q.t == kSolidWall) {
      const double hw = M.eb ? slot_value(q.mf, q.cxf, q.cyf, E).w - E.z : q.hf;
      elem_flux(nx, ny, hw, 0., 0., F[0], F[1], F[2]);
    } else {
      Slot L = slot_value(q.mf, q.cxf, q.cyf, E);
      Slot R = slot_value(q.mt, q.cxt, q.cyt, q.t == kPeriodic ? q.Eq : E);
      if (FLUX == 0) flux_hll<WS>(nx, ny, E.z, L, R, q.edmin, cor, lwi, F, M.eb != 0);
      else flux_hllc<WS>(tx, ty, E.z, L, R, q.edmin, cor, lwi, F, M.eb != 0);
      if (q.neg) { F[0] = -F[0]; F[1] = -F[1]; F[2] = -F[2]; }
    }
    lw = dmin(lw, lwi);
    lwi_item[it] = lwi;
    out_store<NT>(&W.f0[q.e], F[0]);
    out_store<NT>(&W.f1[q.e], F[1]);
    out_store<NT>(&W.f2[q.e], F[2]);
  }
  if (!CFL) return;
  if (!M.cinst) {
    block_min_bits(lw, rbits);
    return;
  }
  // ensembles: one r per instance (4 slots apart); a block inside one
  // instance reduces as usual, a block straddling instances uses atomics
  __shared__ int s_inst, s_mixed;
  if (threadIdx.x == 0) {
    s_inst = in[0].inst;
    s_mixed = 0;
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NI; ++it)
    if (in[it].e < M.e1 && in[it].inst != s_inst) s_mixed = 1;
  __syncthreads();
  if (!s_mixed) {
    block_min_bits(lw, rbits + 4 * (int64_t)s_inst);
  } else {
#pragma unroll
    for (int it = 0; it < NI; ++it)
      if (lwi_item[it] < 1.0)
        atomicMin(rbits + 4 * (int64_t)in[it].inst, (unsigned long long)__double_as_longlong(lwi_item[it]));
  }
}

// ----------------------------------------------------------- draining dt --
// TimeDisc::ComputeDrainingDt (src/TimeDisc.cpp:43-66) on the stage-start
// state.  The reference sums the outflowing F0, a flux per unit length, so
// its "time" A h / sum is not one (D16); with LEN the outflow through each
// edge is F0 * L (the edge length from the cell's own vertices).
template <bool LEN, bool NT>
__global__ void __launch_bounds__(256) k_drain(DevMesh M, DevState S, DevWork W) {
  int64_t i = M.c0 + global_index();
  if (i >= M.c1) return;
  if (M.cinst && !M.bactive[M.cinst[i]]) return;  // finished ensemble member
  double h = S.w[S.at(i)] - M.bc[i];
  double d;
  if (!is_wet(h)) {
    d = 0.;
  } else {
    uint32_t fl = M.cflag[i];
    V3 p0, p1, p2;
    if (LEN) {
      p0 = node(M, M.tp[i]);
      p1 = node(M, M.tp[M.T + i]);
      p2 = node(M, M.tp[2 * M.T + i]);
    }
    double sum = 0.;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double fe = W.f0[M.te[k * M.T + i]];
      double out = sgn_pos(fl, k) ? fe : -fe;
      if (LEN)
        out = out * edge_len(pick(p0, p1, p2, ep0_local(fl, k)), pick(p0, p1, p2, ep1_local(fl, k)));
      sum += dmax(0., out);
    }
    d = sum > kTol ? M.area[i] * h / sum : INFINITY;
  }
  out_store<NT>(&W.dtd[i], d);
}

// -------------------------------------------------------------- update --
// TimeDisc::RHS (src/TimeDisc.cpp:3-41) + stage combination
// (src/Solvers.cpp:11-13,23-25,30-32,42-58) + ConsAssigner.  The inputs of a
// cell are gathered first (RhsIn), the arithmetic follows.
__device__ __forceinline__ void rhs_load(const DevMesh& M, const DevState& S, const DevWork& W,
                                         int i, RhsIn& q) {
  q.fl = M.cflag[i];
  q.mu = load_muscl(S, W, i);
  q.wi = S.w[S.at(i)];
  q.p[0] = node(M, M.tp[i]);
  q.p[1] = node(M, M.tp[M.T + i]);
  q.p[2] = node(M, M.tp[2 * M.T + i]);
  q.dti = W.dtd[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int e = M.te[k * M.T + i];
    const int nb = M.tt[k * M.T + i];
    q.F0[k] = W.f0[e];
    q.F1[k] = W.f1[e];
    q.F2[k] = W.f2[e];
    const double d = W.dtd[nb < 0 ? i : nb];
    q.dtik[k] = nb < 0 ? INFINITY : d;
    // across a periodic edge the neighbour is the partner edge's cell
    if (nb < 0 && M.pe && M.et1[e] == kPeriodic) q.dtik[k] = W.dtd[M.et0[M.pe[e]]];
  }
}
__device__ __forceinline__ void rhs_cell(const DevMesh& M, const DevState& S, const DevWork& W,
                                         int i, double dts, double cor, double fric, double& r0,
                                         double& r1, double& r2) {
  RhsIn q;
  rhs_load(M, S, W, i, q);
  rhs_compute(q, dts, cor, fric, M.eb != 0, r0, r1, r2);
}

__device__ __forceinline__ double stage_dt(const StageArgs& A) {
  double dt = A.dt_fixed;
  if (!(dt > 0.)) dt = A.cfl * __longlong_as_double((long long)A.rbits[0]);
  return dt;
}

// NI cells per thread, BS apart (BS threads per workgroup), loads of all of
// them first (NI and BS are launch-time choices like the flux's).
template <int NI, int BS, bool NT>
__global__ void __launch_bounds__(BS) k_update(DevMesh M, DevState S, DevState U0, DevOut O,
                                                DevWork W, StageArgs A) {
  const int64_t base = M.c0 + (int64_t)logical_block() * (BS * NI) + threadIdx.x;
  const double dt = M.cinst ? 0. : stage_dt(A);
  if (A.dts_log && blockIdx.x == 0 && threadIdx.x == 0) A.dts_log[0] = dt;
  RhsIn in[NI];
  double w[NI], u[NI], v[NI], bc[NI];
  double w0[NI], u0[NI], v0[NI];
  const bool comb = A.a0 != 0.;
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int64_t i = base + it * BS;
    const int is = (int)(i < M.c1 ? i : M.c0);  // out-of-range items read a valid cell and are dropped
    rhs_load(M, S, W, is, in[it]);
    w[it] = S.w[S.at(is)];
    u[it] = S.u[S.at(is)];
    v[it] = S.v[S.at(is)];
    bc[it] = centroid(in[it].p[0], in[it].p[1], in[it].p[2]).z;
    if (comb) {
      w0[it] = U0.w[U0.at(is)];
      u0[it] = U0.u[U0.at(is)];
      v0[it] = U0.v[U0.at(is)];
    }
  }
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int64_t i = base + it * BS;
    if (i >= M.c1) continue;
    double dts = A.coef * dt;
    if (M.cinst) {
      const int inst = M.cinst[i];
      if (!M.bactive[inst]) {
        // finished ensemble member: the stage leaves its state unchanged
        O.w[O.at(i)] = comb ? w0[it] : w[it];
        O.u[O.at(i)] = comb ? u0[it] : u[it];
        O.v[O.at(i)] = comb ? v0[it] : v[it];
        continue;
      }
      dts = A.coef * M.bdt[inst];
    }
    double R0, R1, R2;
    rhs_compute(in[it], dts, A.cor, A.fric, M.eb != 0, R0, R1, R2);
    const double h = w[it] - bc[it];
    const double Uc0 = h, Uc1 = h * u[it], Uc2 = h * v[it];
    double r0, r1, r2;
    if (comb) {
      const double h0 = w0[it] - bc[it];
      const double Ua0 = h0, Ua1 = h0 * u0[it], Ua2 = h0 * v0[it];
      const bool cvx = M.eb && A.a0 + A.a1 == 1.;
      r0 = stage_combine(cvx, A.a0, A.a1, Ua0, Uc0, R0);
      r1 = stage_combine(cvx, A.a0, A.a1, Ua1, Uc1, R1);
      r2 = stage_combine(cvx, A.a0, A.a1, Ua2, Uc2, R2);
    } else {
      r0 = Uc0 + R0;
      r1 = Uc1 + R1;
      r2 = Uc2 + R2;
    }
    double ow, ou, ov;
    cons_assign(bc[it], r0, r1, r2, ow, ou, ov);
    out_store<NT>(&O.w[O.at(i)], ow);
    out_store<NT>(&O.u[O.at(i)], ou);
    out_store<NT>(&O.v[O.at(i)], ov);
  }
}

__global__ void k_reset_r(unsigned long long* rbits) {
  *rbits = (unsigned long long)__double_as_longlong(1.0);
}

// ------------------------------------------------------------ ensembles --
__global__ void k_batch_reset_r(DevBatch B, int slot) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < B.n) B.rbits[4 * k + slot] = (unsigned long long)__double_as_longlong(1.0);
}

// dt of every active member after the step's first flux pass: cfl * r,
// clipped so the last step lands on t_end.
__global__ void k_batch_dt(DevBatch B, double cfl, double t_end) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= B.n) return;
  double d = 0.;
  if (B.active[k]) {
    d = cfl * __longlong_as_double((long long)B.rbits[4 * k]);
    double rest = t_end - B.t[k];
    if (!(d < rest)) d = rest;
  }
  B.dt[k] = d;
}

// end of a step: advance time, retire members that reached t_end, count
// the active ones (one workgroup of 1024 threads, strided over members)
__global__ void __launch_bounds__(1024) k_batch_advance(DevBatch B, double t_end) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int k = threadIdx.x; k < B.n; k += blockDim.x) {
    if (B.active[k]) {
      double d = B.dt[k];
      double t = B.t[k];
      double rest = t_end - t;
      B.t[k] = (d < rest) ? t + d : t_end;
      B.steps[k] += 1;
      B.active[k] = (d < rest) ? 1 : 0;
    }
    mine += B.active[k];
  }
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) *B.n_active = cnt;
}

// ------------------------------------------------------------- getters --
// Edge and source slots as UpdateInterfaceValues leaves them
// (src/SpaceDisc.cpp:15-31), slot 2e + (from < to).
__global__ void __launch_bounds__(256) k_edge_field(DevMesh M, DevState S, DevWork W, double cor,
                                                    double* edg, double* src) {
  int64_t i = global_index();
  if (i >= M.T) return;
  const Muscl mu = load_muscl(S, W, (int)i);
  const double cx = M.cx[i], cy = M.cy[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int e = M.te[k * M.T + i];
    int nb = M.tt[k * M.T + i];
    int64_t s = 2 * (int64_t)e + (((int)i < nb) ? 1 : 0);
    const V3 E = midpoint(node(M, M.ep0[e]), node(M, M.ep1[e]));
    Slot sv = slot_value(mu, cx, cy, E);
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
    double dx = P.x - M.cx[c], dy = P.y - M.cy[c];
    const bool pw = W.cls[c] == 1;
    // pass 1 (a partially wet cell's level is in wrec1) ...
    double a = pw ? W.wrec1[c] + (0. * dx + 0. * dy) : W.mw[c] + (W.mgx[c] * dx + W.mgy[c] * dy);
    if (!((a - P.z) >= 0)) a = P.z;
    mx = dmax(mx, a);
    // ... and pass 2 of partially wet cells
    if (pw) {
      double b = W.mw[c] + (W.mgx[c] * dx + W.mgy[c] * dy);
      if (!((b - P.z) >= 0)) b = P.z;
      mx = dmax(mx, b);
    }
  }
  out[n] = mx;
}

// ---------------------------------------------------------- integrals --
// Per-cell terms of ComputeIntegrals (src/SpaceDisc.cpp:77-103, commented out
// in the reference): area * (h, 1/2 h |u|^2, 1/2 h^2 + h b); partially wet
// cells integrate the energies of their final reconstruction with
// TriangAverage<2,10> (include/PointOperations.h:24-44) at points carried in
// 3-D (the bathymetry at a point is its interpolated z).
__device__ inline void cell_integrals(const DevMesh& M, const DevState& S, const DevWork& W, int i,
                                      double out[3]) {
  out[0] = 0.; out[1] = 0.; out[2] = 0.;
  const int k = W.cls[i];
  if (k == 0) return;
  const double A = M.area[i];
  const double hh = S.w[S.at(i)] - M.bc[i];
  if (k == 2) {
    const double uu = S.u[S.at(i)], vv = S.v[S.at(i)];
    out[0] = A * hh;
    out[1] = A * (0.5 * hh * (uu * uu + vv * vv));
    out[2] = A * (0.5 * hh * hh + hh * M.bc[i]);
    return;
  }
  out[0] = A * hh;
  const double w0 = W.mw[i], gx = W.mgx[i], gy = W.mgy[i], u0 = S.u[S.at(i)], v0 = S.v[S.at(i)];
  const double cx = M.cx[i], cy = M.cy[i];
  const V3 p0 = node(M, M.tp[i]), p1 = node(M, M.tp[M.T + i]), p2 = node(M, M.tp[2 * M.T + i]);
  const int n = 10;
  const double hs = 1. / n;
  const V3 di = {hs * (p1.x - p0.x), hs * (p1.y - p0.y), hs * (p1.z - p0.z)};
  const V3 dj = {hs * (p2.x - p0.x), hs * (p2.y - p0.y), hs * (p2.z - p0.z)};
  const double c = 1. / 3.;
  const V3 dt = {c * (di.x + dj.x), c * (di.y + dj.y), c * (di.z + dj.z)};
  double s0 = 0., s1 = 0.;
  auto acc = [&](const V3& q) {
    const double dx = q.x - cx, dy = q.y - cy;
    double a = w0 + (gx * dx + gy * dy), uq = u0, vq = v0;
    if (!((a - q.z) >= 0)) { a = q.z; uq = 0.; vq = 0.; }
    const double hq = a - q.z;
    s0 += hs * (0.5 * hq * (uq * uq + vq * vq));
    s1 += hs * (0.5 * hq * hq + hq * q.z);
  };
  V3 pi = p0;
  for (int a = 0; a < n; ++a) {
    V3 pt = {pi.x + dt.x, pi.y + dt.y, pi.z + dt.z};
    for (int b = 0; b < n - a - 1; ++b) {
      acc(pt);
      acc({pt.x + dt.x, pt.y + dt.y, pt.z + dt.z});
      pt = {pt.x + dj.x, pt.y + dj.y, pt.z + dj.z};
    }
    acc(pt);
    pi = {pi.x + di.x, pi.y + di.y, pi.z + di.z};
  }
  out[1] = A * (hs * s0);
  out[2] = A * (hs * s1);
}

// fixed-shape tree sum over a 256-thread block (deterministic)
__device__ inline void block_sum3(double v[3], double* red /*3*kBlock*/) {
  const int t = threadIdx.x;
  for (int q = 0; q < 3; ++q) red[q * kBlock + t] = v[q];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (t < off)
      for (int q = 0; q < 3; ++q) red[q * kBlock + t] += red[q * kBlock + t + off];
    __syncthreads();
  }
  for (int q = 0; q < 3; ++q) v[q] = red[q * kBlock];
}

__global__ void __launch_bounds__(256) k_integrals(DevMesh M, DevState S, DevWork W, double* partial,
                                                   double* per_cell) {
  __shared__ double red[3 * kBlock];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double v[3] = {0., 0., 0.};
  if (i < M.T) {
    cell_integrals(M, S, W, (int)i, v);
    if (per_cell)
      for (int q = 0; q < 3; ++q) per_cell[3 * i + q] = v[q];
  }
  block_sum3(v, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) partial[3 * (int64_t)blockIdx.x + q] = v[q];
}

// one block: strided sequential sums of the block partials, then a tree
__global__ void __launch_bounds__(256) k_sum_partials(const double* partial, int64_t nb, double* out) {
  __shared__ double red[3 * kBlock];
  double v[3] = {0., 0., 0.};
  for (int64_t b = threadIdx.x; b < nb; b += kBlock)
    for (int q = 0; q < 3; ++q) v[q] += partial[3 * b + q];
  block_sum3(v, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) out[q] = v[q];
}

// ------------------------------------------------------------ health --
// State statistics for failure detection: non-finite cells, min and max of
// h = w - b, max |u| and max |u| + sqrt(h) over wet cells.  Block partials,
// then one block; min / max and integer counts are order-independent, so
// the result is exact and identical run to run.
struct Stats {
  double nonfinite, hmin, hmax, umax, cmax;
};
__device__ inline void stats_combine(Stats& a, const Stats& b) {
  a.nonfinite += b.nonfinite;
  a.hmin = dmin(a.hmin, b.hmin);
  a.hmax = dmax(a.hmax, b.hmax);
  a.umax = dmax(a.umax, b.umax);
  a.cmax = dmax(a.cmax, b.cmax);
}
__device__ inline Stats block_stats(Stats v) {
  __shared__ Stats red[kBlock];
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (t < off) stats_combine(red[t], red[t + off]);
    __syncthreads();
  }
  return red[0];
}
__global__ void __launch_bounds__(256) k_stats(DevMesh M, DevState S, double* partial) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  Stats v{0., INFINITY, -INFINITY, 0., 0.};
  if (i < M.T) {
    const double w = S.w[S.at(i)], u = S.u[S.at(i)], vv = S.v[S.at(i)];
    if (!isfinite(w) || !isfinite(u) || !isfinite(vv)) {
      v.nonfinite = 1.;
    } else {
      const double h = w - M.bc[i];
      v.hmin = h;
      v.hmax = h;
      if (is_wet(h)) {
        const double sp = sqrt(u * u + vv * vv);
        v.umax = sp;
        v.cmax = sp + sqrt(h);
      }
    }
  }
  const Stats b = block_stats(v);
  if (threadIdx.x == 0) {
    double* p = partial + 5 * (int64_t)blockIdx.x;
    p[0] = b.nonfinite; p[1] = b.hmin; p[2] = b.hmax; p[3] = b.umax; p[4] = b.cmax;
  }
}
__global__ void __launch_bounds__(256) k_stats_final(const double* partial, int64_t nb, double* out) {
  Stats v{0., INFINITY, -INFINITY, 0., 0.};
  for (int64_t b = threadIdx.x; b < nb; b += kBlock) {
    const double* p = partial + 5 * b;
    stats_combine(v, Stats{p[0], p[1], p[2], p[3], p[4]});
  }
  const Stats r = block_stats(v);
  if (threadIdx.x == 0) {
    out[0] = r.nonfinite; out[1] = r.hmin; out[2] = r.hmax; out[3] = r.umax; out[4] = r.cmax;
  }
}

// One reconstruction of one cell for the MUSCLObject::Reconstruct* getters
// (src/MUSCLObject.cpp:31-191): kind 0 dry, 1 PartWet1, 2 FullWet, 3
// PartWet2 (with the pass-1 node maximum of the last interface values).
// out = (origin w, u, v, gradient of w x, y).
template <bool HEAD_GRAD, bool HEAD_PW2>
__global__ void k_reconstruct_one(DevMesh M, DevState S, DevWork W, int i, int kind, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const CellRec ri = M.crec[i];
  const uint32_t fl = M.cflag[i];
  const V3 p0 = node(M, M.tp[i]), p1 = node(M, M.tp[M.T + i]), p2 = node(M, M.tp[2 * M.T + i]);
  const double w = S.w[S.at(i)];
  double o0 = w, o1 = S.u[S.at(i)], o2 = S.v[S.at(i)], gx = 0., gy = 0.;
  const double lvl1 = partwet1_w(ri.zmin, ri.zmin, ri.zmax, ri.bc, w);
  if (kind == 0) {
    o0 = ri.bc; o1 = 0.; o2 = 0.;
    gradient(p0, p1, p2, gx, gy);
  } else if (kind == 1) {
    o0 = lvl1;
  } else if (kind == 2) {
    double wn[3];
    NRec rn[3];
    int jn[3];
    for (int kk = 0; kk < 3; ++kk) {
      int j = M.tt[kk * M.T + i];
      j = j < 0 ? i : j;  // a wall has no neighbour: the cell stands in, as in pass 1
      jn[kk] = j;
      wn[kk] = S.w[S.at(j)];
      const CellRec r = M.crec[j];
      rn[kk] = NRec{r.cx, r.cy, (r.fl & CELL_BND) ? INFINITY : r.zmax, r.bc};
    }
    fullwet_plane<HEAD_GRAD>(M, p0, p1, p2, fl, w, ri.cx, ri.cy, wn, rn, jn, gx, gy);
  } else {
    partwet2_muscl<HEAD_PW2>(M, S, W, i, lvl1, o0, gx, gy);
  }
  out[0] = o0; out[1] = o1; out[2] = o2; out[3] = gx; out[4] = gy;
}

// Fluxes::HLL<W> / HLLC<W> (include/Fluxes.h:14-111) of edge e seen from
// cell `from` (normal outward of it, left state its slot), with the slots
// of the last interface values; `to` < 0 takes the left state for the
// right one.  out = flux (3), then the length-to-wavespeed candidate (or
// +inf when the flux returned before it).
template <int FLUX, int WS>
__global__ void k_edge_flux_one(DevMesh M, DevState S, DevWork W, int e, int from, int to, double cor,
                                double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const V3 Pa = node(M, M.ep0[e]), Pb = node(M, M.ep1[e]);
  const double le = edge_len(Pa, Pb);
  const V3 E = midpoint(Pa, Pb);
  double tx, ty;
  tangent(Pa, Pb, le, V3{M.cx[from], M.cy[from], 0.}, tx, ty);
  const double nx = ty, ny = -tx;
  const Slot L = slot_value(load_muscl(S, W, from), M.cx[from], M.cy[from], E);
  const Slot R = to < 0 ? L : slot_value(load_muscl(S, W, to), M.cx[to], M.cy[to], E);
  const double dl = 2. * M.area[from] / le, dr = to < 0 ? dl : 2. * M.area[to] / le;
  double lw = INFINITY;
  double F[3];
  if (FLUX == 0) flux_hll<WS>(nx, ny, E.z, L, R, dmin(dl, dr), cor, lw, F, M.eb != 0);
  else flux_hllc<WS>(tx, ty, E.z, L, R, dmin(dl, dr), cor, lw, F, M.eb != 0);
  out[0] = F[0]; out[1] = F[1]; out[2] = F[2]; out[3] = lw;
}

__global__ void k_rhs_one(DevMesh M, DevState S, DevWork W, int i, double dt, double cor,
                          double fric, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  rhs_cell(M, S, W, i, dt, cor, fric, out[0], out[1], out[2]);
}

// ------------------------------------------------------ static scratch --
// Block partials of the allocation-free integrals (swe_integrals_noalloc):
// a file-scope array, one copy per GPU, for meshes of up to kScratchBlocks
// workgroups of 256 cells.  Host code never names it in a launch: it takes
// the array's device address from scratch_address().
__device__ __align__(256) double scratch[3 * kScratchBlocks];

hipError_t scratch_address(double** out) {
  static std::mutex mu;
  static std::vector<double*> addr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  if (addr.size() <= size_t(dev)) addr.resize(size_t(dev) + 1, nullptr);
  if (!addr[size_t(dev)]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[size_t(dev)] = static_cast<double*>(p);
  }
  *out = addr[size_t(dev)];
  return hipSuccess;
}

// -------------------------------------------------------------- launch --

hipError_t launch_setup_geometry(const DevMesh& M, const DevGeomOut& G, hipStream_t st) {
  hipLaunchKernelGGL(k_setup_nodes, grid_for(M.V), dim3(256), 0, st, M, G);
  hipLaunchKernelGGL(k_setup_cells, grid_for(M.T), dim3(256), 0, st, M, G);
  hipLaunchKernelGGL(k_setup_edges, grid_for(M.E), dim3(256), 0, st, M, G);
  return hipGetLastError();
}
hipError_t launch_recon1(const DevMesh& M, const DevState& S, const DevWork& W, int opts,
                         hipStream_t st, bool zero_count) {
  if (zero_count) hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, st, W.pw_count);
  const dim3 g = grid_for(M.c1 - M.c0);
  if (opts & 1) {
    if (M.nt) hipLaunchKernelGGL((k_recon1<true, true>), g, dim3(256), 0, st, M, S, W);
    else hipLaunchKernelGGL((k_recon1<true, false>), g, dim3(256), 0, st, M, S, W);
  } else {
    if (M.nt) hipLaunchKernelGGL((k_recon1<false, true>), g, dim3(256), 0, st, M, S, W);
    else hipLaunchKernelGGL((k_recon1<false, false>), g, dim3(256), 0, st, M, S, W);
  }
  return hipGetLastError();
}
hipError_t launch_recon2(const DevMesh& M, const DevState& S, const DevWork& W, int opts,
                         hipStream_t st) {
  dim3 g = grid_for(M.c1 - M.c0);
  if (g.x > 1024) g.x = 1024;  // grid-stride over the list
  if (opts & 2) hipLaunchKernelGGL((k_recon2<true>), g, dim3(256), 0, st, M, S, W);
  else hipLaunchKernelGGL((k_recon2<false>), g, dim3(256), 0, st, M, S, W);
  return hipGetLastError();
}
__global__ void __launch_bounds__(256) k_relayout(DevState S, DevOut O, int64_t T) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= T) return;
  const int64_t a = S.at(i), b = O.at(i);
  O.w[b] = S.w[a];
  O.u[b] = S.u[a];
  O.v[b] = S.v[a];
}
hipError_t launch_relayout(const DevState& S, const DevOut& O, int64_t T, hipStream_t st) {
  if (T > 0) hipLaunchKernelGGL(k_relayout, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, st, S, O, T);
  return hipGetLastError();
}
hipError_t launch_flux(const DevMesh& M, const DevState& S, const DevWork& W, int flux, int ws,
                       double cor, unsigned long long* rbits, int items, int block, hipStream_t st,
                       bool reset) {
  // without the reset (a fixed-dt pass of a single mesh) the minimum is not
  // taken at all; ensembles always take their per-member minima
  const bool cfl = reset || M.cinst;
  if (reset && !M.cinst) hipLaunchKernelGGL(k_reset_r, dim3(1), dim3(1), 0, st, rbits);
  if (M.e1 <= M.e0) return hipGetLastError();
  if (block != 256 && block != 512) return hipErrorInvalidValue;
  dim3 g = grid_for_bs((M.e1 - M.e0 + items - 1) / items, block), b(block);
#define SWE_FLUX_LAUNCH(F, Wv, N, BSV)                                                                       \
  if (cfl && M.nt) hipLaunchKernelGGL((k_flux<F, Wv, N, true, BSV, true>), g, b, 0, st, M, S, W, cor, rbits);  \
  else if (cfl) hipLaunchKernelGGL((k_flux<F, Wv, N, true, BSV, false>), g, b, 0, st, M, S, W, cor, rbits);   \
  else if (M.nt) hipLaunchKernelGGL((k_flux<F, Wv, N, false, BSV, true>), g, b, 0, st, M, S, W, cor, rbits); \
  else hipLaunchKernelGGL((k_flux<F, Wv, N, false, BSV, false>), g, b, 0, st, M, S, W, cor, rbits);
#define SWE_FLUX_CASE(F, Wv, N)                                                        \
  if (flux == F && ws == Wv && items == N) {                                           \
    if (block == 256) { SWE_FLUX_LAUNCH(F, Wv, N, 256) }                               \
    else { SWE_FLUX_LAUNCH(F, Wv, N, 512) }                                            \
    return hipGetLastError();                                                          \
  }
#define SWE_FLUX_CASES(N)                                                              \
  SWE_FLUX_CASE(0, 0, N) SWE_FLUX_CASE(0, 1, N) SWE_FLUX_CASE(0, 2, N)                 \
  SWE_FLUX_CASE(1, 0, N) SWE_FLUX_CASE(1, 1, N) SWE_FLUX_CASE(1, 2, N)
  SWE_FLUX_CASES(1) SWE_FLUX_CASES(2) SWE_FLUX_CASES(4)
#undef SWE_FLUX_CASES
#undef SWE_FLUX_CASE
#undef SWE_FLUX_LAUNCH
  return hipErrorInvalidValue;
}
hipError_t launch_drain(const DevMesh& M, const DevState& S, const DevWork& W, int opts,
                        hipStream_t st) {
  const dim3 g = grid_for(M.c1 - M.c0);
  const bool len = (opts & kOptDrainLength) != 0;
  if (len && M.nt) hipLaunchKernelGGL((k_drain<true, true>), g, dim3(256), 0, st, M, S, W);
  else if (len) hipLaunchKernelGGL((k_drain<true, false>), g, dim3(256), 0, st, M, S, W);
  else if (M.nt) hipLaunchKernelGGL((k_drain<false, true>), g, dim3(256), 0, st, M, S, W);
  else hipLaunchKernelGGL((k_drain<false, false>), g, dim3(256), 0, st, M, S, W);
  return hipGetLastError();
}
hipError_t launch_update(const DevMesh& M, const DevState& S, const DevState& U0, const DevOut& O,
                         const DevWork& W, const StageArgs& A, int items, int block, hipStream_t st) {
  if (block != 256 && block != 512) return hipErrorInvalidValue;
  dim3 g = grid_for_bs((M.c1 - M.c0 + items - 1) / items, block), b(block);
#define SWE_UPD(N, BSV)                                                                              \
  if (items == N && block == BSV) {                                                                  \
    if (M.nt) hipLaunchKernelGGL((k_update<N, BSV, true>), g, b, 0, st, M, S, U0, O, W, A);           \
    else hipLaunchKernelGGL((k_update<N, BSV, false>), g, b, 0, st, M, S, U0, O, W, A);               \
    return hipGetLastError();                                                                        \
  }
  SWE_UPD(1, 256) SWE_UPD(2, 256) SWE_UPD(1, 512) SWE_UPD(2, 512)
#undef SWE_UPD
  return hipErrorInvalidValue;
}
hipError_t launch_batch_reset_r(const DevBatch& B, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_reset_r, dim3((B.n + 255) / 256), dim3(256), 0, st, B, slot);
  return hipGetLastError();
}
hipError_t launch_batch_dt(const DevBatch& B, double cfl, double t_end, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_dt, dim3((B.n + 255) / 256), dim3(256), 0, st, B, cfl, t_end);
  return hipGetLastError();
}
hipError_t launch_batch_advance(const DevBatch& B, double t_end, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(1024), 0, st, B, t_end);
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
int64_t integrals_blocks(const DevMesh& M) { return (M.T + kBlock - 1) / kBlock; }
hipError_t launch_stats(const DevMesh& M, const DevState& S, double* partial, double* out,
                        hipStream_t st) {
  const int64_t nb = integrals_blocks(M);
  if (nb > 0) hipLaunchKernelGGL(k_stats, dim3((unsigned)nb), dim3(256), 0, st, M, S, partial);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(256), 0, st, partial, nb, out);
  return hipGetLastError();
}
hipError_t launch_integrals(const DevMesh& M, const DevState& S, const DevWork& W, double* partial,
                            double* per_cell, double* out, hipStream_t st) {
  const int64_t nb = integrals_blocks(M);
  if (nb > 0)
    hipLaunchKernelGGL(k_integrals, dim3((unsigned)nb), dim3(256), 0, st, M, S, W, partial, per_cell);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial, nb, out);
  return hipGetLastError();
}
hipError_t launch_reconstruct_one(const DevMesh& M, const DevState& S, const DevWork& W, int i, int kind,
                                  int opts, double* out, hipStream_t st) {
  const bool hg = (opts & 1) != 0, hp = (opts & 2) != 0;
  if (hg && hp) hipLaunchKernelGGL((k_reconstruct_one<true, true>), dim3(1), dim3(64), 0, st, M, S, W, i, kind, out);
  else if (hg) hipLaunchKernelGGL((k_reconstruct_one<true, false>), dim3(1), dim3(64), 0, st, M, S, W, i, kind, out);
  else if (hp) hipLaunchKernelGGL((k_reconstruct_one<false, true>), dim3(1), dim3(64), 0, st, M, S, W, i, kind, out);
  else hipLaunchKernelGGL((k_reconstruct_one<false, false>), dim3(1), dim3(64), 0, st, M, S, W, i, kind, out);
  return hipGetLastError();
}
hipError_t launch_edge_flux_one(const DevMesh& M, const DevState& S, const DevWork& W, int flux, int ws, int e,
                                int from, int to, double cor, double* out, hipStream_t st) {
#define SWE_EF(F, Wv) \
  if (flux == F && ws == Wv) { hipLaunchKernelGGL((k_edge_flux_one<F, Wv>), dim3(1), dim3(64), 0, st, M, S, W, e, from, to, cor, out); return hipGetLastError(); }
  SWE_EF(0, 0) SWE_EF(0, 1) SWE_EF(0, 2) SWE_EF(1, 0) SWE_EF(1, 1) SWE_EF(1, 2)
#undef SWE_EF
  return hipErrorInvalidValue;
}
hipError_t launch_rhs_one(const DevMesh& M, const DevState& S, const DevWork& W, int i, double dt,
                          double cor, double fric, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_rhs_one, dim3(1), dim3(64), 0, st, M, S, W, i, dt, cor, fric, out);
  return hipGetLastError();
}

}  // namespace swe
