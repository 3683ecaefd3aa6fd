// Fused second half of a stage: ComputeFluxes, the draining time steps and
// the update (src/SpaceDisc.cpp:54-74, src/TimeDisc.cpp:3-66,
// src/Solvers.cpp:5-59) in one kernel per tile of kTileCells cells.
//
// A workgroup evaluates the flux of every edge of its tile's cells and of
// the cells around the tile (the ring) once, into LDS; then the draining
// time step of every local cell from those fluxes; then the update of its
// own cells.  Fluxes and draining steps never go to HBM: the separate flux,
// draining and update kernels write 24 B per edge and 8 B per cell and read
// them back by gathers.  Edges between two tiles are evaluated by both
// (each tile needs them), and so are the ring cells' other edges; every
// evaluation and every update is the same sequence of operations as in the
// separate kernels (the same device functions), so the results are
// identical bit for bit.
//
// The stage step size must be known before the kernel: a fixed dt, or the
// CFL step of an earlier flux pass of the same step (stages 2 and 3).  The
// pass's own CFL length-to-wavespeed minimum is not taken: the solver
// recomputes it with the fluxes when a getter asks for them.
//
// The first stage of a CFL step learns its step size from its own flux pass,
// so the separate flux kernel runs first; the FF instance of this kernel
// then loads the listed edges' fluxes from memory instead of evaluating them
// (the same values: the separate flux kernel is the same device functions
// in the same orientation) and runs the draining and update phases as
// usual, in place of the separate draining and update kernels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "swe_common.h"
#include "swe_tiles.h"

namespace swe {

__device__ __forceinline__ double stage_dt_fused(const StageArgs& A) {
  double dt = A.dt_fixed;
  if (!(dt > 0.)) dt = A.cfl * __longlong_as_double((long long)A.rbits[0]);
  return dt;
}

// waves per SIMD the register allocation must allow, and threads per tile
// workgroup (kTileCells or 2 kTileCells: the flux and draining phases then
// give each thread fewer items, the update leaves the second half idle);
// build-time knobs
#ifndef SWE_FUSED_MIN_WAVES
#define SWE_FUSED_MIN_WAVES 4
#endif
#ifndef SWE_FUSED_THREADS
#define SWE_FUSED_THREADS 256
#endif
constexpr int kFusedThreads = SWE_FUSED_THREADS;
static_assert(kFusedThreads == kTileCells || kFusedThreads == 2 * kTileCells, "threads per tile");

template <int FLUX, int WS, bool LEN, bool NT, bool FF>
__global__ void __launch_bounds__(kFusedThreads, SWE_FUSED_MIN_WAVES) k_fused(DevMesh M, DevTiles P, DevState S, DevState U0,
                                                      DevOut O, DevWork W, StageArgs A, int64_t t0, int64_t t1,
                                                      int store) {
  __shared__ double sF0[kTileMaxEval], sF1[kTileMaxEval], sF2[kTileMaxEval];
  __shared__ double sD[kTileMaxLocal];
  // the evaluating instance walks the tiles backwards (it follows a
  // forward reconstruction); the FF instance forwards (it follows k_flux)
  const int64_t tile = t0 + (FF ? logical_block() : logical_block_reversed());
  if (tile >= t1) return;  // padding workgroups of the XCD-aware grid
  const int64_t c0 = tile * kTileCells;
  const int nc = (int)(M.T - c0 < kTileCells ? M.T - c0 : kTileCells);
  const int rg0 = P.ring_off[tile], nr = P.ring_off[tile + 1] - rg0;
  const int e0 = P.ev_off[tile], ne = P.ev_off[tile + 1] - e0;
  const int tid = threadIdx.x;
  const int64_t T = M.T;

  // ---- edge fluxes (k_flux's operations), one evaluation per LDS slot;
  // the edge record gives both cells and nodes (one load level, as the
  // separate flux kernel's edge tables)
  // index loads of the later phases, issued before the flux phase: the
  // own cell's slot codes, flags and vertex ids, and the thread's ring
  // entry (nl <= kTileMaxLocal = 2 kTileCells <= kTileCells + kFusedThreads:
  // at most one per thread)
  const bool mine = tid < nc;
  const int i = (int)(c0 + (mine ? tid : 0));
  uint32_t code[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) code[k] = P.code[k * T + i];
  const CVtx cvi = M.cv[i];
  const int tpi[3] = {cvi.v0, cvi.v1, cvi.v2};
  const uint32_t fli = cvi.fl;
  const int nl = kTileCells + nr;
  const int lr = kTileCells + tid;
  int xr = 0;
  uint32_t cdr[3] = {0u, 0u, 0u};
  if (lr < nl) {
    xr = P.ring_cell[rg0 + tid];
#pragma unroll
    for (int k = 0; k < 3; ++k) cdr[k] = P.ring_code[3 * (rg0 + tid) + k];
  }

  // the own cell's update and draining inputs, and the ring cell's
  // draining inputs (loaded before the flux phase by the FF instance, whose
  // flux phase is loads only; after it by the evaluating one, whose flux
  // phase needs the registers)
  RhsCell q;
  double w, u, v, w0 = 0., u0 = 0., v0 = 0.;
  const bool comb = A.a0 != 0.;
  auto load_own = [&]() {
    q.fl = fli;
    q.mu = load_muscl(S, W, i);
    q.wi = S.w[S.at(i)];
    q.p[0] = node(M, tpi[0]);
    q.p[1] = node(M, tpi[1]);
    q.p[2] = node(M, tpi[2]);
    w = S.w[S.at(i)];
    u = S.u[S.at(i)];
    v = S.v[S.at(i)];
    if (comb) {
      w0 = U0.w[U0.at(i)];
      u0 = U0.u[U0.at(i)];
      v0 = U0.v[U0.at(i)];
    }
  };
  double hr = 0., arear = 0.;
  uint32_t flr = 0u;
  V3 rp0{0., 0., 0.}, rp1{0., 0., 0.}, rp2{0., 0., 0.};
  auto load_ring = [&]() {
    if (lr < nl) {
      const CVtx cvr = M.cv[xr];
      if (LEN) {
        rp0 = node(M, cvr.v0);
        rp1 = node(M, cvr.v1);
        rp2 = node(M, cvr.v2);
      }
      hr = S.w[S.at(xr)] - M.bc[xr];
      flr = cvr.fl;
      arear = M.area[xr];
    }
  };
  if constexpr (FF) {
    load_own();
    load_ring();
    // fluxes of the stage's flux pass, written by the separate flux kernel;
    // ne <= kTileMaxEval: at most kFE evaluations per thread, every load of
    // them issued before the first LDS write
    constexpr int kFE = (kTileMaxEval + kFusedThreads - 1) / kFusedThreads;
    int ev[kFE];
#pragma unroll
    for (int it = 0; it < kFE; ++it) {
      const int j = tid + it * kFusedThreads;
      ev[it] = j < ne ? P.evedge[e0 + j] : -1;
    }
    double f0[kFE], f1[kFE], f2[kFE];
#pragma unroll
    for (int it = 0; it < kFE; ++it) {
      const int e = ev[it] < 0 ? 0 : ev[it];
      f0[it] = W.f0[e];
      f1[it] = W.f1[e];
      f2[it] = W.f2[e];
    }
#pragma unroll
    for (int it = 0; it < kFE; ++it) {
      const int j = tid + it * kFusedThreads;
      if (j < ne) {
        sF0[j] = f0[it];
        sF1[j] = f1[it];
        sF2[j] = f2[it];
      }
    }
  } else {
    // the next evaluation's edge record is loaded before this one's gathers
    EdgeRec rn = tid < ne ? P.evrec[e0 + tid] : EdgeRec{0, 0, 0, 0};
    for (int j = tid; j < ne; j += kFusedThreads) {
      const EdgeRec r = rn;
      if (j + kFusedThreads < ne) rn = P.evrec[e0 + j + kFusedThreads];
      const int f = r.et0, t = r.et1;
      const int tr = t < 0 ? f : t;  // boundary edges read their own cell twice: R = L
      const V3 Pa = node(M, r.ep0), Pb = node(M, r.ep1);
      const Muscl mf = load_muscl(S, W, f), mt = load_muscl(S, W, tr);
      const CXY cf = M.cxy[f], ct = M.cxy[tr];
      const double cxf = cf.x, cyf = cf.y, cxt = ct.x, cyt = ct.y;
      const double hf = (t == kSolidWall && !M.eb) ? S.w[S.at(f)] - M.bc[f] : 0.;
      const double le = edge_len(Pa, Pb);
      const V3 E = midpoint(Pa, Pb);
      double tx, ty;
      tangent(Pa, Pb, le, V3{cxf, cyf, 0.}, tx, ty);
      const double nx = ty, ny = -tx;
      double F[3];
      double lw = 1.0;  // the CFL candidate is not taken here (unused)
      if (t == kSolidWall) {
        // SWE_OPT_EXACT_BALANCE: the inside edge depth instead of the cell mean
        const double hw = M.eb ? slot_value(mf, cxf, cyf, E).w - E.z : hf;
        elem_flux(nx, ny, hw, 0., 0., F[0], F[1], F[2]);
      } else {
        const Slot L = slot_value(mf, cxf, cyf, E);
        const Slot R = slot_value(mt, cxt, cyt, E);
        if (FLUX == 0) flux_hll<WS>(nx, ny, E.z, L, R, 1.0, A.cor, lw, F, M.eb != 0);
        else flux_hllc<WS>(tx, ty, E.z, L, R, 1.0, A.cor, lw, F, M.eb != 0);
      }
      sF0[j] = F[0];
      sF1[j] = F[1];
      sF2[j] = F[2];
      if (store && f >= c0 && f < c0 + nc) {  // the edge's one owning evaluation: its et0 cell is ours
        const int e = P.evedge[e0 + j];
        out_store<NT>(&W.f0[e], F[0]);
        out_store<NT>(&W.f1[e], F[1]);
        out_store<NT>(&W.f2[e], F[2]);
      }
    }
  }
  __syncthreads();

  // ---- draining time steps of every local cell (k_drain's operations).
  // The thread's own cell (tid < nc) is also the cell it updates: its
  // update inputs are loaded here with its draining inputs, so one latency
  // covers both phases' loads.
  if constexpr (!FF) load_own();
  // the draining step of a cell from its depth, edges and flux slots
  auto drain_dt = [&](double h, uint32_t fl, const uint32_t* cd, const V3& p0, const V3& p1, const V3& p2,
                      double area) -> double {
    if (!is_wet(h)) return 0.;
    double sum = 0.;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double fe = sF0[cd[k] >> 16];
      double out = sgn_pos(fl, k) ? fe : -fe;
      if (LEN) out = out * edge_len(pick(p0, p1, p2, ep0_local(fl, k)), pick(p0, p1, p2, ep1_local(fl, k)));
      sum += dmax(0., out);
    }
    return sum > kTol ? area * h / sum : INFINITY;
  };
  if (mine) {
    // the centroid bathymetry and area from the vertices, as the setup
    // kernel computes M.bc and M.area (the same functions: the same bits)
    const double d = drain_dt(w - centroid(q.p[0], q.p[1], q.p[2]).z, q.fl, code, q.p[0], q.p[1], q.p[2],
                              tri_area(q.p[0], q.p[1], q.p[2]));
    sD[tid] = d;
    if (store) out_store<NT>(&W.dtd[i], d);
  }
  if constexpr (!FF) load_ring();
  if (lr < nl) sD[lr] = drain_dt(hr, flr, cdr, rp0, rp1, rp2, arear);
  __syncthreads();

  // ---- update of the tile's cells (k_update's operations; the fluxes and
  // neighbour draining steps are read from LDS edge by edge)
  if (!mine) return;
  const double dt = M.cinst ? 0. : stage_dt_fused(A);
  if (A.dts_log && tile == 0 && tid == 0) A.dts_log[0] = dt;
  double dts = A.coef * dt;
  if (M.cinst) {  // ensembles: the member's own step size (k_update's rule)
    const int inst = M.cinst[i];
    if (!M.bactive[inst]) {
      // finished member, or a cell of a tile that straddles the edge of the
      // active window: the stage leaves its state unchanged
      const int64_t o = O.at(i);
      O.w[o] = comb ? w0 : w;
      O.u[o] = comb ? u0 : u;
      O.v[o] = comb ? v0 : v;
      return;
    }
    dts = A.coef * M.bdt[inst];
  }
  q.dti = sD[tid];
  const double bc = centroid(q.p[0], q.p[1], q.p[2]).z;
  double R0, R1, R2;
  rhs_compute_by(
      q,
      [&](int k, double& F0, double& F1, double& F2, double& dtik) {
        const int s = (int)(code[k] >> 16);
        const uint32_t nb = code[k] & 0xffffu;
        F0 = sF0[s];
        F1 = sF1[s];
        F2 = sF2[s];
        dtik = nb >= kNbOut ? INFINITY : sD[nb];
      },
      dts, A.cor, A.fric, M.eb != 0, R0, R1, R2);
  const double hh = w - bc;
  const double Uc0 = hh, Uc1 = hh * u, Uc2 = hh * v;
  double r0, r1, r2;
  if (comb) {
    const double h0 = w0 - bc;
    const double Ua0 = h0, Ua1 = h0 * u0, Ua2 = h0 * v0;
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
  cons_assign(bc, r0, r1, r2, ow, ou, ov);
  out_store<NT>(&O.w[O.at(i)], ow);
  out_store<NT>(&O.u[O.at(i)], ou);
  out_store<NT>(&O.v[O.at(i)], ov);
}

hipError_t launch_fused(const DevMesh& M, const DevTiles& P, const DevState& S, const DevState& U0,
                        const DevOut& O, const DevWork& W, const StageArgs& A, int flux, int ws, int opts,
                        hipStream_t st, bool store, bool from_fluxes) {
  // the tiles that hold cells of the visited window [c0, c1): the whole mesh,
  // or an ensemble's running members (cells of finished members in the
  // first and last tile are copied unchanged, as every finished member's)
  if (!M.cinst && (M.c0 != 0 || M.c1 != M.T)) return hipErrorInvalidValue;
  const int64_t t0 = M.c0 / kTileCells, t1 = (M.c1 + kTileCells - 1) / kTileCells;
  if (t1 <= t0) return hipGetLastError();
  const dim3 g = grid_for_bs(t1 - t0, 1), b(kFusedThreads);
  const bool len = (opts & kOptDrainLength) != 0;
  const int sto = store ? 1 : 0;
  if (from_fluxes) {  // the flux scheme does not enter: one instance per drain / store variant
    if (len && M.nt) hipLaunchKernelGGL((k_fused<0, 0, true, true, true>), g, b, 0, st, M, P, S, U0, O, W, A, t0, t1, sto);
    else if (len) hipLaunchKernelGGL((k_fused<0, 0, true, false, true>), g, b, 0, st, M, P, S, U0, O, W, A, t0, t1, sto);
    else if (M.nt) hipLaunchKernelGGL((k_fused<0, 0, false, true, true>), g, b, 0, st, M, P, S, U0, O, W, A, t0, t1, sto);
    else hipLaunchKernelGGL((k_fused<0, 0, false, false, true>), g, b, 0, st, M, P, S, U0, O, W, A, t0, t1, sto);
    return hipGetLastError();
  }
#define SWE_FUSED(F, Wv)                                                                             \
  if (flux == F && ws == Wv) {                                                                       \
    if (len && M.nt) hipLaunchKernelGGL((k_fused<F, Wv, true, true, false>), g, b, 0, st, M, P, S, U0, O, W, A, t0, t1, sto);   \
    else if (len) hipLaunchKernelGGL((k_fused<F, Wv, true, false, false>), g, b, 0, st, M, P, S, U0, O, W, A, t0, t1, sto);     \
    else if (M.nt) hipLaunchKernelGGL((k_fused<F, Wv, false, true, false>), g, b, 0, st, M, P, S, U0, O, W, A, t0, t1, sto);    \
    else hipLaunchKernelGGL((k_fused<F, Wv, false, false, false>), g, b, 0, st, M, P, S, U0, O, W, A, t0, t1, sto);             \
    return hipGetLastError();                                                                        \
  }
  SWE_FUSED(0, 0) SWE_FUSED(0, 1) SWE_FUSED(0, 2) SWE_FUSED(1, 0) SWE_FUSED(1, 1) SWE_FUSED(1, 2)
#undef SWE_FUSED
  return hipErrorInvalidValue;
}

}  // namespace swe
