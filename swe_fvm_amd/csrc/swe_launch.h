// Device data layout and kernel launchers of the explicit step.
//
// HBM layout (structure of arrays, int32 indices, Morton-ordered cells):
//   nodes  px, py, pz                     double[V], and pn {x, y, z, 0} (32 B)
//   cells  tp, te, tt                     int32[3][T]   (column k at k*T)
//          cflag                          uint32[T]
//            bits 0-2  et0(te[k]) == cell (flux orientation of edge k)
//            bit  3    boundary triangle (some edge has et1 < 0)
//            bits 4+4k, 6+4k  local vertex of ep0 / ep1 of edge k
//   edges  ep0, ep1, et0, et1             int32[E], pe (periodic partner) int32[E]
//   node -> cell map  n2c_off[V+1], n2c[3T]
//   static geometry (computed once by the setup kernels with the step's own
//   formulas, so every use is bit-identical to recomputing it):
//     cells  cx, cy, bc (centroid x, y, z), area, and an AoS copy
//            crec {cx, cy, bc, zmin, zmax, cflag} (48 B) and its 32-byte
//            part nrec {cx, cy, zmax or +inf on boundary cells, bc} for the
//            reconstruction's own-cell and neighbour reads, and cxy (cx, cy
//            interleaved) for the flux evaluations
//     edges  edmin (min over the two cells of 2 area / length; the inside
//            cell alone on a free-flow edge, 0 on a wall)
//   Edge lengths, midpoints and tangents are recomputed from the nodes.
//   state  w, u, v                        double[T] per buffer
//   work   MUSCL (mw, mgx, mgy): pass 1, overwritten in place by pass 2 for
//          partially wet cells, whose pass-1 level is kept in wrec1 for the
//          node-maximum readers; cls int8; fluxes f0, f1, f2 double[E];
//          draining dt dtd double[T]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

// static record of a cell for the reconstruction's neighbour gathers: one
// 48-byte load instead of six scattered ones
struct __attribute__((aligned(16))) CellRec {
  double cx, cy, bc, zmin, zmax;
  uint32_t fl, pad;
};

// the part of a neighbour's record the full-wet reconstruction reads in its
// common branch: centroid, highest vertex (+inf on boundary cells, which
// never count as full-wet neighbours) and centroid bathymetry; 32 bytes,
// staged in LDS by k_recon1
struct __attribute__((aligned(32))) NRec {
  double cx, cy, zmx, bc;
};

// a cell's centroid (x, y) in one 16-byte load, for the flux evaluations'
// gathers of both cells of an edge
struct __attribute__((aligned(16))) CXY {
  double x, y;
};

// a node's coordinates (x, y, z) in one aligned 32-byte record
struct __attribute__((aligned(32))) NodeRec {
  double x, y, z, pad;
};

// a cell's vertex ids and flags (tp[0..2][i], cflag[i]) in one 16-byte load
struct __attribute__((aligned(16))) CVtx {
  int v0, v1, v2;
  uint32_t fl;
};

struct DevMesh {
  const double* px;
  const double* py;
  const double* pz;
  const int* tp;
  const int* te;
  const int* tt;
  const uint32_t* cflag;
  const int* ep0;
  const int* ep1;
  const int* et0;
  const int* et1;
  const int* n2c_off;
  const int* n2c;
  int64_t T, E, V;
  // static geometry
  const double* cx;
  const double* cy;
  const double* bc;
  const double* area;
  const double* edmin;
  const CellRec* crec;
  // batched ensembles (null for a single mesh): instance of every cell and
  // the per-instance step size and activity flag
  const int* cinst;
  const double* bdt;
  const int* bactive;
  // cells [c0, c1) and edges [e0, e1) the step kernels visit (the whole mesh
  // except when finished ensemble members are trimmed off the front)
  int64_t c0, c1, e0, e1;
  // outputs that only later kernels read are stored nontemporally (meshes of
  // 2^20 cells or more: a stage's data exceeds the 256 MB Infinity Cache;
  // smaller ones keep them cached for the next kernel)
  int nt;
  // partner edge of each PERIODIC edge (et1 == -3), -1 elsewhere; null when
  // the mesh has none
  const int* pe;
  // SWE_OPT_EXACT_BALANCE: the flux and RHS groupings that keep equal edge
  // states at rest to the bit
  int eb;
  // the 32-byte part of crec the reconstruction reads for every cell and
  // neighbour (highest vertex +inf on boundary cells); crec itself is read
  // only where zmin or the true zmax of a boundary cell enter
  const NRec* nrec;
  // (cx, cy) interleaved
  const CXY* cxy;
  // (px, py, pz) interleaved, kept equal to them by the setup kernels
  const NodeRec* pn;
  // (tp[0][i], tp[1][i], tp[2][i], cflag[i]) interleaved
  const CVtx* cv;
};

// per-instance control of a batched ensemble
struct DevBatch {
  int n;                      // instances
  double* t;                  // simulated time
  double* dt;                 // step size of the current step
  int* active;                // t < t_end
  long long* steps;           // steps taken
  unsigned long long* rbits;  // r of each flux pass, 4 slots per instance
  int* n_active;              // number of active instances after the last step
};

// writable view of the static geometry arrays, for the setup kernel
struct DevGeomOut {
  double *cx, *cy, *bc, *area, *edmin;
  CellRec* crec;
  NRec* nrec;
  CXY* cxy;
  NodeRec* pn;
  CVtx* cv;
};

// Cell-state layout: the w, u, v of cell i sit at w[at(i)], u[at(i)],
// v[at(i)] with at(i) = 3 i - 2 (i & mask).  The default mask (2^31 - 1)
// gives at(i) = i: three separate arrays (SoA).  mask = B - 1 for a power of
// two B gives blocks of B cells holding B values of w, then of u, then of v
// (AoSoA; B = 1 is AoS), with u and v the pointers B and 2 B past w.
constexpr uint32_t kLayoutSoA = 0x7fffffffu;
constexpr int kLayoutMaxBlock = 64;  // largest block; allocations are padded to a multiple of it
__host__ __device__ __forceinline__ int64_t state_at(int64_t i, uint32_t mask) {
  return 3 * i - 2 * (int64_t)((uint32_t)i & mask);
}
struct DevState {
  const double* w;
  const double* u;
  const double* v;
  uint32_t mask = kLayoutSoA;
  __host__ __device__ __forceinline__ int64_t at(int64_t i) const { return state_at(i, mask); }
};

struct DevOut {
  double* w;
  double* u;
  double* v;
  uint32_t mask = kLayoutSoA;
  __host__ __device__ __forceinline__ int64_t at(int64_t i) const { return state_at(i, mask); }
};

struct DevWork {
  double* mw;
  double* mgx;
  double* mgy;
  double* wrec1;  // PartWet1 level of partially wet cells (pass 1)
  int8_t* cls;
  double* f0;
  double* f1;
  double* f2;
  double* dtd;
  int* pw_list;            // partially wet cells of the current pass 1
  unsigned int* pw_count;  // their number
};

// adjoint work of one stage (reverse kernels, swe_adjoint.hip); layouts:
// per cell [q][T] or [k][T] (q component, k local edge), per edge side
// [q][side][E] (side 0: et0, 1: et1)
struct DevAdj {
  const double* obar;  // [3][T] adjoint of the stage output (w, u, v)
  double* sbar;        // [3][T] adjoint of the stage input S
  double* u0bar;       // [3][T] adjoint of the step's first state U0
  double* rb;          // [3][T] adjoint of the cell's conserved update r
  double* dtd_own;     // [T] draining-dt adjoint from the cell's own update
  double* dtd_nb;      // [3][T] ... of neighbour k, from this cell's update
  double* m_cell;      // [5][T] MUSCL (w0, gx, gy) adjoint the own update leaves
  double* sl_edge;     // [3][2][E] adjoint of the edge slot (w, u, v) per side;
                       // wall edges: side 0 w holds the et0 cell's depth adjoint
  double* drain_h;     // [T] depth adjoint of the draining dt (k_adj_flux)
  double* w_nb;        // [3][T] w adjoint of neighbour k from this cell's plane
  double* zc;          // [3][T] node-z adjoint per cell corner, summed over stages
  int8_t* back;        // [3][T] local index of this cell in neighbour k
  int8_t* eloc;        // [2][E] local index of edge e in its et0 / et1 cell
  // partially wet cells (k_adj_pw, k_adj_pwnode): their own w and corner-z
  // adjoints, the adjoint of the node maximum at their lowest vertex, the
  // latter routed to the cell attaining the maximum (per corner of that
  // cell, cleared by its consumer) or, summed over stages, to the node's
  // own height; the height adjoint of a neighbour's vertex opposite an edge
  // (partially wet neighbour levels), per edge side, summed over stages
  double* pw_own;      // [4][T] w, corner z 0..2
  double* pw1_jac;     // [4][T] d PartWet1 level / d (w, corner z 0..2)
  double* pw_w23bar;   // [T]
  double* p1bar;       // [3][T] routed pass-1 level adjoint per corner
  double* zn_acc;      // [V]
  double* zopp_acc;    // [2][E]
};

// tangents of the tangent-linear step (swe_tangent.hip): the node
// bathymetry's (input) and the stage intermediates' (written per stage)
struct DevTan {
  const double* zd;    // [V] node z tangent
  double* mw;          // [T] MUSCL level, gradient tangents
  double* mgx;
  double* mgy;
  double* wrec1;       // [T] pass-1 level tangent of partially wet cells
  double* f0;          // [E] flux tangents
  double* f1;
  double* f2;
  double* dtd;         // [T] draining-dt tangent
};

struct StageArgs {
  double a0, a1;       // Shu-Osher weights of U0 and the stage state (a0 = 0: U + RHS)
  double coef;         // stage dt = coef * dt
  double dt_fixed;     // > 0 fixed step, else cfl * r of the step's first flux pass
  double cfl;
  double cor;
  const unsigned long long* rbits;  // r of the step's first flux pass (double bits)
  double* dts_log;     // receives the step's dt (first stage only), may be null
  double fric;         // linear friction coefficient (0: none)
};

// the state of cells [0, T) rewritten from S's layout into O's
hipError_t launch_relayout(const DevState& S, const DevOut& O, int64_t T, hipStream_t st);
hipError_t launch_setup_geometry(const DevMesh& M, const DevGeomOut& G, hipStream_t st);
// zero_count false: the caller has zeroed W.pw_count (launch_slot_counts)
hipError_t launch_recon1(const DevMesh& M, const DevState& S, const DevWork& W, int opts,
                         hipStream_t st, bool zero_count = true);
hipError_t launch_recon2(const DevMesh& M, const DevState& S, const DevWork& W, int opts,
                         hipStream_t st);
// items: edges per thread of the flux (1, 2 or 4), cells per thread of the
// update (1 or 2); block: threads per workgroup of either (256 or 512);
// results are identical for every choice
// reset false (single mesh): no CFL minimum at all, neither reset nor taken
// (a fixed-dt pass that never reads it); the fluxes are the same
hipError_t launch_flux(const DevMesh& M, const DevState& S, const DevWork& W, int flux, int ws,
                       double cor, unsigned long long* rbits, int items, int block, hipStream_t st,
                       bool reset = true);
// opts & 32 (SWE_OPT_DRAIN_LENGTH): outflow F0 * L in the draining dt
hipError_t launch_drain(const DevMesh& M, const DevState& S, const DevWork& W, int opts,
                        hipStream_t st);
hipError_t launch_update(const DevMesh& M, const DevState& S, const DevState& U0, const DevOut& O,
                         const DevWork& W, const StageArgs& A, int items, int block, hipStream_t st);
hipError_t launch_batch_reset_r(const DevBatch& B, int slot, hipStream_t st);
hipError_t launch_batch_dt(const DevBatch& B, double cfl, double t_end, hipStream_t st);
hipError_t launch_batch_advance(const DevBatch& B, double t_end, hipStream_t st);
hipError_t launch_edge_field(const DevMesh& M, const DevState& S, const DevWork& W, double cor,
                             double* edg, double* src, hipStream_t st);
hipError_t launch_max_wp(const DevMesh& M, const DevWork& W, double* out, hipStream_t st);
// mass, kinetic and potential energy of the current reconstruction: block
// partials (integrals_blocks(M) x 3 doubles) reduced in a fixed order into
// out[3]; per_cell (optional, 3 T doubles) receives the cell terms
int64_t integrals_blocks(const DevMesh& M);
hipError_t launch_integrals(const DevMesh& M, const DevState& S, const DevWork& W, double* partial,
                            double* per_cell, double* out, hipStream_t st);
// state statistics (non-finite count, min h, max h, max |u|, max |u| +
// sqrt(h)): partial needs 5 * integrals_blocks(M) doubles, out 5
hipError_t launch_stats(const DevMesh& M, const DevState& S, double* partial, double* out,
                        hipStream_t st);
// reverse of one stage: from J.obar (whose w part still owes the previous
// reverse stage's neighbour terms wnb_in, when not null) to J.sbar (whose w
// part owes J.w_nb), accumulating the node-z adjoint per cell corner in
// J.zc.  u0mode: kU0Accumulate adds this stage's U0 adjoint to J.u0bar
// (else it is set, for stages that combine U0), kU0Into adds J.u0bar into
// J.sbar (the first stage, whose input is U0).  The forward intermediates of
// the stage state S must be in W (recon1, recon2, flux, drain of S).
// npw: the length of W's pass-2 list when the caller knows it (the
// partially wet kernels are then skipped for an empty list and sized to it
// otherwise), kPwUnknown when not.
constexpr int kU0Accumulate = 1, kU0Into = 2;
constexpr unsigned int kPwUnknown = 0xffffffffu;
hipError_t launch_adjoint_stage(const DevMesh& M, const DevState& S, const DevState& U0, const DevWork& W,
                                const StageArgs& A, const DevAdj& J, int flux, int ws, int opts,
                                const double* wnb_in, int u0mode, unsigned int npw, hipStream_t st);
// the pass-2 list lengths of n stage slots whose counters lie stride bytes
// apart from first: zeroed (out null) or copied to out[n]
hipError_t launch_slot_counts(unsigned int* first, int64_t stride, int n, unsigned int* out, hipStream_t st);
// one stage of the tangent-linear step: values (S, U0 -> O, intermediates
// in W) exactly as the forward stage, tangents (Sd, U0d, Q.zd -> Od,
// intermediates in Q); fixed dt (A.dt_fixed)
hipError_t launch_tangent_stage(const DevMesh& M, const DevState& S, const DevState& Sd, const DevState& U0,
                                const DevState& U0d, const DevOut& O, const DevOut& Od, const DevWork& W,
                                const DevTan& Q, const StageArgs& A, int flux, int ws, int opts,
                                hipStream_t st);
// the mesh tables J.back and J.eloc (once per mesh)
hipError_t launch_adjoint_tables(const DevMesh& M, const DevAdj& J, hipStream_t st);
// after the last reverse stage: J.sbar's w part gathers J.w_nb; zbar (if not
// null) += the corners of J.zc at every node
hipError_t launch_adjoint_finish(const DevMesh& M, const DevAdj& J, double* zbar, hipStream_t st);
// Fluxes, draining time steps and update of one stage in one kernel over
// tiles (swe_fused.hip); the reconstruction of S must be in W.  The stage's
// step size must be known (A.dt_fixed > 0, or the CFL r of an earlier pass
// in A.rbits[0], or an ensemble's per-member steps).  W's fluxes and
// draining time steps are written only with store (each edge's flux and each
// cell's draining step once); this pass's CFL minimum never is.  With
// from_fluxes the stage's fluxes are already in W (the separate flux pass
// of a CFL step's first stage) and are read instead of evaluated.
struct DevTiles;
hipError_t launch_fused(const DevMesh& M, const DevTiles& P, const DevState& S, const DevState& U0,
                        const DevOut& O, const DevWork& W, const StageArgs& A, int flux, int ws, int opts,
                        hipStream_t st, bool store = false, bool from_fluxes = false);
// The allocation-free integrals keep their block partials in a static
// per-GPU array of 3 * kScratchBlocks doubles (meshes up to kScratchBlocks *
// 256 cells); scratch_address gives its device address on the current
// device (looked up once per device).
constexpr int64_t kScratchBlocks = 65536;
hipError_t scratch_address(double** out);
// one reconstruction (kind 0 dry, 1 PartWet1, 2 FullWet, 3 PartWet2) of
// cell i: out[5] = origin (w, u, v), gradient of w (x, y)
hipError_t launch_reconstruct_one(const DevMesh& M, const DevState& S, const DevWork& W, int i, int kind,
                                  int opts, double* out, hipStream_t st);
// flux of edge e from cell `from` to cell `to` (< 0: right state = left):
// out[4] = flux, length-to-wavespeed candidate (+inf if none)
hipError_t launch_edge_flux_one(const DevMesh& M, const DevState& S, const DevWork& W, int flux, int ws, int e,
                                int from, int to, double cor, double* out, hipStream_t st);
hipError_t launch_rhs_one(const DevMesh& M, const DevState& S, const DevWork& W, int i, double dt,
                          double cor, double fric, double* out, hipStream_t st);

}  // namespace swe
