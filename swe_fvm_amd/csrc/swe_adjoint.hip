// Reverse (adjoint) kernels of one explicit stage.
//
// A stage maps the stage state S and the step's first state U0 (primitive
// w, u, v per cell) to its output O (src/Solvers.cpp:5-59 with
// TimeDisc::RHS, src/TimeDisc.cpp:3-66).  Given the adjoint of O, these
// kernels return the adjoints of S and U0 and add the adjoint of the node
// bathymetry z.  The forward intermediates (MUSCL, fluxes, draining dt) are
// those the forward kernels wrote for this stage, so the linearisation point
// is bit-identical to the forward step.  Branch
// decisions (wet/dry, the limiter's all-or-nothing choices, min / max
// selections) are the forward ones; the Riemann flux is differentiated by
// its reverse mode (swe_flux_adj.h), the transpose of the flux template's
// dual-number derivative.
//
// Every contribution is written to a slot owned by one (cell, side) or
// (edge, side) and gathered in a fixed order by its consumer, so the
// reverse pass is atomic-free and identical run to run, like the forward.
//
// Partially wet cells: the local Jacobians of PartWet1 and PartWet2 come
// from their generic restatements (swe_pw.h) on dual numbers; the adjoint of
// PartWet2's node maximum goes to the cell attaining it (k_adj_pw,
// k_adj_pwnode), so the whole reconstruction is differentiated.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "swe_common.h"
#include "swe_flux_adj.h"
#include "swe_pw.h"

namespace swe {

// local index of edge e in cell c (te[k][c] == e)
__device__ __forceinline__ int local_edge(const DevMesh& M, int c, int e) {
  return M.te[c] == e ? 0 : (M.te[M.T + c] == e ? 1 : 2);
}
// local index k' with tt[k'][j] == c
__device__ __forceinline__ int back_index(const DevMesh& M, int j, int c) {
  return M.tt[j] == c ? 0 : (M.tt[M.T + j] == c ? 1 : 2);
}

// sum over the neighbours j of cell c of their per-side value toward c
__device__ __forceinline__ double gather_nb(const DevMesh& M, const int8_t* back, const double* wnb, int c) {
  const int64_t T = M.T;
  double s = 0.;
  for (int k = 0; k < 3; ++k) {
    const int j = M.tt[k * T + c];
    if (j >= 0) s += wnb[back[k * T + c] * T + j];
  }
  return s;
}

// value and attaining cell (-1: the node's own height) of the pass-1 node
// maximum at node n (node_max_pass1)
__device__ inline double node_max_arg(const DevMesh& M, const DevWork& W, int n, const V3& P, int& arg) {
  double mx = P.z;
  arg = -1;
  for (int q = M.n2c_off[n]; q < M.n2c_off[n + 1]; ++q) {
    const int c = M.n2c[q];
    double w0 = W.mw[c], gx = W.mgx[c], gy = W.mgy[c];
    if (W.cls[c] == 1) { w0 = W.wrec1[c]; gx = 0.; gy = 0.; }
    const double dx = P.x - M.cx[c], dy = P.y - M.cy[c];
    double a = w0 + (gx * dx + gy * dy);
    if (!((a - P.z) >= 0)) a = P.z;
    if (mx < a) {
      mx = a;
      arg = c;
    }
  }
  return mx;
}

// lowest vertex of cell c by PartWet2's sort (recon2_cell), and the highest
// vertex height
__device__ inline int lowest_node(const DevMesh& M, int c, double& zmx) {
  const int64_t T = M.T;
  int i0 = M.tp[c], i1 = M.tp[T + c], i2 = M.tp[2 * T + c];
  double z0 = M.pz[i0], z1 = M.pz[i1], z2 = M.pz[i2];
  double tz;
  int ti;
  if (z0 > z1) { tz = z0; z0 = z1; z1 = tz; ti = i0; i0 = i1; i1 = ti; }
  if (z1 > z2) { tz = z1; z1 = z2; z2 = tz; ti = i1; i1 = i2; i2 = ti; }
  if (z0 > z1) { ti = i0; i0 = i1; i1 = ti; }
  zmx = z2;
  return i0;
}

// Final MUSCL (w0, gx, gy) of partially wet cell c: PartWet2, or the
// PartWet1 level where pass 2 falls back, as functions of (w, z of local
// vertices 0..2, pass-1 node maximum at the lowest vertex); true when the
// node maximum enters.  p1[4]: the PartWet1 level's derivatives in (w, z of
// local vertices 0..2), the cell's pass-1 value where a node maximum or a
// full-wet neighbour's plane takes it.
__device__ inline bool pw2_jac(const DevMesh& M, const DevState& S, const DevWork& W, int c, double jac[3][5],
                               double p1[4]) {
  using D = Dual<5>;
  const int64_t T = M.T;
  int i0 = M.tp[c], i1 = M.tp[T + c], i2 = M.tp[2 * T + c];
  PT<D> Q0{dconst<5>(M.px[i0]), dconst<5>(M.py[i0]), dvar<5>(M.pz[i0], 1)};
  PT<D> Q1{dconst<5>(M.px[i1]), dconst<5>(M.py[i1]), dvar<5>(M.pz[i1], 2)};
  PT<D> Q2{dconst<5>(M.px[i2]), dconst<5>(M.py[i2]), dvar<5>(M.pz[i2], 3)};
  const double t = 1. / 3.;
  const D bc = (Q0.z * t + Q1.z * t) + Q2.z * t;
  const D w = dvar<5>(S.w[c], 0);
  const D zmn = dmin(dmin(Q0.z, Q1.z), Q2.z), zmx = dmax(dmax(Q0.z, Q1.z), Q2.z);
  D w0 = partwet1_g(zmn, zmn, zmx, bc, w), gx = dconst<5>(0.), gy = dconst<5>(0.);
  for (int k = 0; k < 4; ++k) p1[k] = w0.d[k];
  bool used = false;
  sort3_g(Q0, Q1, Q2, i0, i1, i2);
  if (!((w.v > Q2.z.v) || (Q2.z.v - Q0.z.v < kTol))) {
    int arg;
    const double w23 = node_max_arg(M, W, i0, V3{Q0.x.v, Q0.y.v, Q0.z.v}, arg);
    D a, bx, by;
    if (partwet2_g(Q0, Q1, Q2, M.cx[c], M.cy[c], bc, w, dvar<5>(w23, 4), a, bx, by)) {
      w0 = a;
      gx = bx;
      gy = by;
      used = true;
    }
  }
  for (int k = 0; k < 5; ++k) {
    jac[0][k] = w0.d[k];
    jac[1][k] = gx.d[k];
    jac[2][k] = gy.d[k];
  }
  return used;
}

// Reverse of slot_value (AtPoint + PrimAssigner, src/SpaceDisc.cpp:24-25):
// from the adjoint (sw, su, sv) of the slot to the MUSCL adjoint mb[5]
// (w0, gx, gy, u0, v0) and the edge-midpoint z adjoint ez.
__device__ inline void slot_bar(const Muscl& m, double cx, double cy, const V3& E, double sw,
                                double su, double sv, double mb[5], double& ez) {
  const double dx = E.x - cx, dy = E.y - cy;
  const double a = m.w0 + (m.gx * dx + m.gy * dy);
  const bool wet = (a - E.z) >= 0;
  if (!wet) {  // AtPoint returns (z, 0, 0); PrimAssigner keeps it
    ez += sw;
    return;
  }
  const double h = a - E.z;
  if (!is_wet(h)) {
    ez += sw;
    return;
  }
  double abar = sw, hbar = 0.;
  if (h < 1e-3) {
    const double q = h * h + 1e-6;
    const double fac = sqrt(2.) * h / sqrt(q);
    const double dfac = sqrt(2.) * 1e-6 / (q * sqrt(q));
    mb[3] += fac * su;
    mb[4] += fac * sv;
    hbar += (m.u0 * su + m.v0 * sv) * dfac;
  } else {
    mb[3] += su;
    mb[4] += sv;
  }
  abar += hbar;
  ez -= hbar;
  mb[0] += abar;
  mb[1] += abar * dx;
  mb[2] += abar * dy;
}

// --------------------------------------------------------------- update --
// Reverse of k_update for cell i: ConsAssigner, the stage combination and
// TimeDisc::RHS.  Writes the adjoint of the conserved update (from which
// k_adj_flux weighs each edge side), the draining-dt adjoints (own,
// neighbours'), the MUSCL adjoint, the U0 adjoint, the first part of the S
// adjoint and the corner z adjoint.
template <bool NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_adj_update(DevMesh M, DevState S, DevState U0, DevWork W,
                                                    StageArgs A, DevAdj J, const double* wnb_in,
                                                    int u0mode) {
  const int64_t i64 = global_index();  // XCD-aware, as the forward kernels
  if (i64 >= M.T) return;
  const int i = (int)i64;
  const int64_t T = M.T;
  const CVtx cvi = M.cv[i];
  const uint32_t fl = cvi.fl;
  const int cls = W.cls[i];  // loads come before the first store (see below)
  const Muscl mu = load_muscl(S, W, i);
  const V3 p[3] = {node(M, cvi.v0), node(M, cvi.v1), node(M, cvi.v2)};
  const V3 c = centroid(p[0], p[1], p[2]);
  const double bc = c.z;
  const double i_area = 1. / tri_area(p[0], p[1], p[2]);
  const double dts = A.coef * A.dt_fixed;
  const double dti = W.dtd[i];
  const double w = S.w[i], u = S.u[i], v = S.v[i];
  const bool comb = A.a0 != 0.;
  // forward recomputation of R (rhs_compute) keeping the per-edge terms
  double F[3][3], dtk[3], dtik[3], hk[3], src1[3], src2[3], cek[3], nxk[3], nyk[3];
  V3 E[3];
  int ek[3];
  double R0 = 0., R1 = 0., R2 = 0.;
  for (int k = 0; k < 3; ++k) {
    const int e = M.te[k * T + i];
    const int nb = M.tt[k * T + i];
    ek[k] = e;
    F[k][0] = W.f0[e]; F[k][1] = W.f1[e]; F[k][2] = W.f2[e];
    dtik[k] = nb < 0 ? INFINITY : W.dtd[nb];
    const bool own = sgn_pos(fl, k);
    const double sg = own ? 1. : -1.;
    dtk[k] = (sg * F[k][0]) > 0. ? dmin(dts, dti) : dmin(dts, dtik[k]);
    const V3 Pa = pick(p[0], p[1], p[2], ep0_local(fl, k));
    const V3 Pb = pick(p[0], p[1], p[2], ep1_local(fl, k));
    const double le = edge_len(Pa, Pb);
    cek[k] = i_area * le;
    E[k] = midpoint(Pa, Pb);
    const Slot s = slot_value(mu, c.x, c.y, E[k]);
    hk[k] = s.w - E[k].z;
    src1[k] = mu.gx + A.cor * (-s.v);
    src2[k] = mu.gy + A.cor * s.u;
    double tx, ty;
    tangent(Pa, Pb, le, c, tx, ty);
    nxk[k] = ty;
    nyk[k] = -tx;
    const double cf = dtk[k] * sg * cek[k];
    R0 -= cf * F[k][0];
    R1 -= cf * F[k][1];
    R2 -= cf * F[k][2];
    const double cs = dts * (1. / 3.);
    R1 -= cs * src1[k] * hk[k];
    R2 -= cs * src2[k] * hk[k];
    const double q2 = 0.5 * hk[k] * hk[k];
    R1 += dtk[k] * (nxk[k] * cek[k] * q2);
    R2 += dtk[k] * (nyk[k] * cek[k] * q2);
  }
  const double hh = w - bc;
  if (A.fric != 0.) {
    R1 -= dts * (A.fric * (hh * mu.u0));
    R2 -= dts * (A.fric * (hh * mu.v0));
  }
  const double Uc0 = hh, Uc1 = hh * u, Uc2 = hh * v;
  const double w0p = U0.w[i], u0p = U0.u[i], v0p = U0.v[i];
  const double h0 = w0p - bc;
  double r0, r1, r2;
  if (comb) {
    r0 = (A.a0 * h0 + A.a1 * Uc0) + R0;
    r1 = (A.a0 * (h0 * u0p) + A.a1 * Uc1) + R1;
    r2 = (A.a0 * (h0 * v0p) + A.a1 * Uc2) + R2;
  } else {
    r0 = Uc0 + R0;
    r1 = Uc1 + R1;
    r2 = Uc2 + R2;
  }
  // ConsAssigner reverse (src/Assigners.cpp:28-44)
  // the w adjoint still owes the previous reverse stage's neighbour terms
  double ow = J.obar[i];
  if (wnb_in) ow += gather_nb(M, J.back, wnb_in, i);
  const double ou = J.obar[T + i], ov = J.obar[2 * T + i];
  double rb0 = 0., rb1 = 0., rb2 = 0., bcb = ow;
  if (is_wet(r0)) {
    double ih, dih;
    if (r0 < 1e-3) {
      const double q = r0 * r0 * r0 * r0 + 1e-12;
      ih = sqrt(2.) * r0 / sqrt(q);
      dih = sqrt(2.) * (1e-12 - r0 * r0 * r0 * r0) / (q * sqrt(q));
    } else {
      ih = 1. / r0;
      dih = -ih * ih;
    }
    rb0 = ow + dih * (r1 * ou + r2 * ov);
    rb1 = ih * ou;
    rb2 = ih * ov;
  }
  // the corner z adjoint accumulated over the stages, and the first stage's
  // U0 adjoint, read before this kernel's first store: the compiler may not
  // move a load of one buffer past a store to another
  double zc_in[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) zc_in[k] = J.zc[k * T + i];
  const bool into_early = (u0mode & kU0Into) && !comb;
  double u0in[3] = {0., 0., 0.};
  if (into_early)
#pragma unroll
    for (int k = 0; k < 3; ++k) u0in[k] = J.u0bar[k * T + i];
  // stage combination
  double sw = 0., su = 0., sv = 0.;
  const double ac = comb ? A.a1 : 1.;
  {
    const double g0 = ac * rb0, g1 = ac * rb1, g2 = ac * rb2;  // adjoint of (h, hu, hv) of S
    const double hb = g0 + u * g1 + v * g2;
    sw += hb;
    bcb -= hb;
    su += hh * g1;
    sv += hh * g2;
  }
  if (comb) {  // the U0 adjoint, summed over the stages that combine U0
    const double g0 = A.a0 * rb0, g1 = A.a0 * rb1, g2 = A.a0 * rb2;
    const double hb = g0 + u0p * g1 + v0p * g2;
    if (u0mode & kU0Accumulate) {
      J.u0bar[i] += hb;
      J.u0bar[T + i] += h0 * g1;
      J.u0bar[2 * T + i] += h0 * g2;
    } else {
      J.u0bar[i] = hb;
      J.u0bar[T + i] = h0 * g1;
      J.u0bar[2 * T + i] = h0 * g2;
    }
    bcb -= hb;
  }
  if (into_early) {  // first stage: its input is U0 itself
    sw += u0in[0];
    su += u0in[1];
    sv += u0in[2];
  } else if (u0mode & kU0Into) {
    sw += J.u0bar[i];
    su += J.u0bar[T + i];
    sv += J.u0bar[2 * T + i];
  }
  // TimeDisc::RHS reverse
  double mb[5] = {0., 0., 0., 0., 0.};
  double dtib = 0.;
  if (A.fric != 0.) {
    const double hhb = -dts * A.fric * (mu.u0 * rb1 + mu.v0 * rb2);
    mb[3] += -dts * A.fric * hh * rb1;
    mb[4] += -dts * A.fric * hh * rb2;
    sw += hhb;
    bcb -= hhb;
  }
  double zc[3] = {0., 0., 0.};
  // outputs read only by later kernels: out_store (nontemporal on large meshes)
  out_store<NT>(&J.rb[i], rb0);  // the flux adjoint of side k is -dtk sg cek rb (k_adj_flux)
  out_store<NT>(&J.rb[T + i], rb1);
  out_store<NT>(&J.rb[2 * T + i], rb2);
  for (int k = 0; k < 3; ++k) {
    const bool own = sgn_pos(fl, k);
    const double sg = own ? 1. : -1.;
    const double q2 = 0.5 * hk[k] * hk[k];
    double dtkb = -(F[k][0] * rb0 + F[k][1] * rb1 + F[k][2] * rb2) * sg * cek[k];
    dtkb += nxk[k] * cek[k] * q2 * rb1 + nyk[k] * cek[k] * q2 * rb2;
    const double cs = dts * (1. / 3.);
    double hkb = dtk[k] * (nxk[k] * cek[k] * rb1 + nyk[k] * cek[k] * rb2) * hk[k];
    hkb += -cs * (src1[k] * rb1 + src2[k] * rb2);
    const double s1b = -cs * hk[k] * rb1, s2b = -cs * hk[k] * rb2;
    mb[1] += s1b;
    mb[2] += s2b;
    const double svb = -A.cor * s1b, sub = A.cor * s2b;
    double ez = -hkb;
    slot_bar(mu, c.x, c.y, E[k], hkb, sub, svb, mb, ez);
    const int a0 = ep0_local(fl, k), a1 = ep1_local(fl, k);
    zc[a0] += 0.5 * ez;
    zc[a1] += 0.5 * ez;
    // draining dt selection (dmin(dts, d) picks d when d < dts)
    // (written only where drain_bar reads them: draining dt below dts)
    double nbb = 0.;
    if ((sg * F[k][0]) > 0.) {
      if (dti < dts) dtib += dtkb;
    } else {
      if (dtik[k] < dts) nbb = dtkb;
    }
    if (dtik[k] < dts) out_store<NT>(&J.dtd_nb[k * T + i], nbb);
  }
  (void)ek;
  if (dti < dts) out_store<NT>(&J.dtd_own[i], dtib);
  // the update's own MUSCL adjoint, applied here where the reconstruction's
  // reverse is a fixed map: the level to the cell's w (full-wet cells, and
  // partially wet ones above every vertex, which pass 2 leaves at w) or to
  // the corners through b_cell (dry cells); u0, v0 to the cell's u, v (zero
  // for dry cells).  What stays for k_adj_recon / k_adj_pw: the gradient
  // part of dry and full-wet cells, all of it for cells in pass 2's list.
  const bool listed = cls == 1 && !(w > dmax(dmax(p[0].z, p[1].z), p[2].z));
  if (cls == 0) {
    bcb += mb[0];
  } else {
    if (!listed) sw += mb[0];
    su += mb[3];
    sv += mb[4];
  }
  if (listed) out_store<NT>(&J.m_cell[i], mb[0]);
  if (cls != 1 || listed) {
    out_store<NT>(&J.m_cell[T + i], mb[1]);
    out_store<NT>(&J.m_cell[2 * T + i], mb[2]);
  }
  const double cthird = 1. / 3.;
  for (int k = 0; k < 3; ++k)  // over every stage
    out_store<NT>(&J.zc[k * T + i], zc_in[k] + (zc[k] + cthird * bcb));
  out_store<NT>(&J.sbar[i], sw);
  out_store<NT>(&J.sbar[T + i], su);
  out_store<NT>(&J.sbar[2 * T + i], sv);
}

// ---------------------------------------------------------------- drain --
// Reverse of k_drain for cell c: the draining dt adjoint of the cell (its
// own update's and its neighbours') into the mass fluxes of its edges (sb per
// unit outflow, on the edges with out[k] > 0) and its depth (hb).  The dt
// adjoint is nonzero only where the draining dt was below the stage dt, so
// the consumers call this for those cells only; false when nothing flows.
template <bool LEN>
__device__ inline bool drain_bar(const DevMesh& M, const DevState& S, const DevWork& W, const DevAdj& J,
                                 int c, double& sb, double& hb, double out[3], double len[3]) {
  const int64_t T = M.T;
  double db = J.dtd_own[c];
  for (int k = 0; k < 3; ++k) {
    const int j = M.tt[k * T + c];
    if (j >= 0) db += J.dtd_nb[J.back[k * T + c] * T + j];
  }
  if (db == 0.) return false;
  const double h = S.w[c] - M.bc[c];
  if (!is_wet(h)) return false;
  const uint32_t fl = M.cflag[c];
  V3 p0, p1, p2;
  if (LEN) {
    p0 = node(M, M.tp[c]);
    p1 = node(M, M.tp[T + c]);
    p2 = node(M, M.tp[2 * T + c]);
  }
  double sum = 0.;
  for (int k = 0; k < 3; ++k) {
    const double fe = W.f0[M.te[k * T + c]];
    out[k] = sgn_pos(fl, k) ? fe : -fe;
    len[k] = 1.;
    if (LEN) {
      len[k] = edge_len(pick(p0, p1, p2, ep0_local(fl, k)), pick(p0, p1, p2, ep1_local(fl, k)));
      out[k] = out[k] * len[k];
    }
    sum += dmax(0., out[k]);
  }
  if (!(sum > kTol)) return false;
  const double A = M.area[c];
  hb = db * A / sum;
  sb = -db * A * h / (sum * sum);
  return true;
}

// The flux adjoint of cell c's side of edge e (local index k, sign sg, the
// cell across o or < 0): -dtk sg cek rb, as TimeDisc::RHS weighs the edge's
// flux in the cell's update, plus the draining term on the mass flux when
// the cell's draining dt was active; the side of the cell's local edge 0
// also stores the draining depth term for k_adj_recon.
template <bool LEN, bool NT>
__device__ inline void side_flux_bar(const DevMesh& M, const DevState& S, const DevWork& W, const DevAdj& J,
                                     int c, int o, int k, double sg, double f0, double le, double dts,
                                     double fb[3]) {
  const int64_t T = M.T;
  const double dtc = W.dtd[c];
  const double dto = o < 0 ? INFINITY : W.dtd[o];
  const double dtk = (sg * f0) > 0. ? dmin(dts, dtc) : dmin(dts, dto);
  const double cek = (1. / M.area[c]) * le;
  const double cf = dtk * sg * cek;
  fb[0] = -cf * J.rb[c];
  fb[1] = -cf * J.rb[T + c];
  fb[2] = -cf * J.rb[2 * T + c];
  if (dtc < dts) {
    double sb, hb, out[3], len[3];
    const bool act = drain_bar<LEN>(M, S, W, J, c, sb, hb, out, len);
    if (k == 0) out_store<NT>(&J.drain_h[c], act ? hb : 0.);  // once per cell, from its local edge 0
    if (act && out[k] > 0.) fb[0] += sb * len[k] * sg;
  }
}

// ----------------------------------------------------------------- flux --
// Reverse of k_flux for edge e: the flux adjoint of its two sides (from the
// cells' update adjoints), through the reverse mode of the Riemann solver
// (inputs L (w, u, v), R (w, u, v); edge z by the chain rule), into the
// adjoints of the two slot values; k_adj_recon takes them through the slot
// reconstruction.
template <int FLUX, int WS, bool LEN, bool NT>
__global__ void __launch_bounds__(256) k_adj_flux(DevMesh M, DevState S, DevWork W, DevAdj J, double dts) {
  // XCD-aware as the forward kernels, each range walked backwards: the
  // update adjoint before it and the reconstruction adjoint after it walk
  // forwards, so each kernel starts on the lines the previous touched last
  const int64_t e64 = (int64_t)logical_block_reversed() * kBlock + threadIdx.x;
  if (e64 >= M.E) return;
  const int e = (int)e64;
  const int64_t Ecnt = M.E;
  const int f = M.et0[e], t = M.et1[e];
  const V3 Pa = node(M, M.ep0[e]), Pb = node(M, M.ep1[e]);
  const double le = edge_len(Pa, Pb);
  const double f0 = W.f0[e];
  double Fb[3], fb[3];
  side_flux_bar<LEN, NT>(M, S, W, J, f, t, J.eloc[e], 1., f0, le, dts, Fb);
  if (t >= 0) {
    side_flux_bar<LEN, NT>(M, S, W, J, t, f, J.eloc[Ecnt + e], -1., f0, le, dts, fb);
    for (int q = 0; q < 3; ++q) Fb[q] += fb[q];
  }
  double g[7] = {0., 0., 0., 0., 0., 0., 0.};
  if (t == kSolidWall) {
    // ElemFlux(n, (h, 0, 0)) = (0, h^2/2 n) of the cell average: the depth
    // adjoint travels in side 0's w slot
    const CXY cf = M.cxy[f];
    double tx, ty;
    tangent(Pa, Pb, le, V3{cf.x, cf.y, 0.}, tx, ty);
    const double nx = ty, ny = -tx;
    const double hf = S.w[f] - M.bc[f];
    if (is_wet(hf)) g[0] = (Fb[1] * nx + Fb[2] * ny) * hf;
  } else {
    const int tr = t < 0 ? f : t;
    const Muscl mf = load_muscl(S, W, f), mt = load_muscl(S, W, tr);
    const CXY cf = M.cxy[f], ct = M.cxy[tr];
    const V3 E = midpoint(Pa, Pb);
    double tx, ty;
    tangent(Pa, Pb, le, V3{cf.x, cf.y, 0.}, tx, ty);
    const double nx = ty, ny = -tx;
    const Slot L = slot_value(mf, cf.x, cf.y, E);
    const Slot R = slot_value(mt, ct.x, ct.y, E);
    // reverse mode of the Riemann solver (swe_flux_adj.h); the edge
    // bathymetry enters it only through hl = L.w - z and hr = R.w - z
    // (include/Fluxes.h:18-21,63-66): dF/dz = -(dF/dL.w + dF/dR.w)
    if (FLUX == 0) flux_hll_vjp<WS>(nx, ny, E.z, L, R, Fb, g);
    else flux_hllc_vjp<WS>(tx, ty, E.z, L, R, Fb, g);
    g[6] = -(g[0] + g[3]);
    if (t < 0) {  // free flow: R is L
      for (int k = 0; k < 3; ++k) {
        g[k] += g[3 + k];
        g[3 + k] = 0.;
      }
    }
  }
  // the edge z term g[6] = -(g[0] + g[3]) is formed by k_adj_recon from the
  // two w slots (free flow: g[3] is merged into g[0], the sum unchanged)
  for (int q = 0; q < 3; ++q) {
    out_store<NT>(&J.sl_edge[(q * 2 + 0) * Ecnt + e], g[q]);
    out_store<NT>(&J.sl_edge[(q * 2 + 1) * Ecnt + e], g[3 + q]);
  }
}

// ------------------------------------------------------ partially wet --
// For each partially wet cell (pass 2's list): its final MUSCL adjoint (its
// own update's and its edge slots', gathered as k_adj_recon gathers it)
// through PartWet2 / PartWet1: the cell's own w and corner-z adjoints, and
// the adjoint of the node maximum at its lowest vertex; and the Jacobian of
// its PartWet1 level, which k_adj_recon applies where that level enters a
// node maximum or a full-wet neighbour's plane (no dual numbers there).
__global__ void __launch_bounds__(256) k_adj_pw(DevMesh M, DevState S, DevWork W, DevAdj J) {
  const unsigned npw = *W.pw_count;
  const int64_t T = M.T, Ecnt = M.E;
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < npw; q += gridDim.x * blockDim.x) {
    const int i = W.pw_list[q];
    const uint32_t fl = M.cflag[i];
    double mb[5] = {J.m_cell[i], J.m_cell[T + i], J.m_cell[2 * T + i], 0., 0.};  // u0, v0: unused here
    const Muscl mu = load_muscl(S, W, i);
    const V3 p0 = node(M, M.tp[i]), p1 = node(M, M.tp[T + i]), p2 = node(M, M.tp[2 * T + i]);
    for (int k = 0; k < 3; ++k) {
      const int e = M.te[k * T + i];
      if (M.et1[e] == kSolidWall) continue;
      const int side = sgn_pos(fl, k) ? 0 : 1;
      double ez = 0.;
      slot_bar(mu, M.cx[i], M.cy[i], edge_mid(p0, p1, p2, fl, k), J.sl_edge[(0 * 2 + side) * Ecnt + e],
               J.sl_edge[(1 * 2 + side) * Ecnt + e], J.sl_edge[(2 * 2 + side) * Ecnt + e], mb, ez);
    }
    double jac[3][5], lj[4];
    const bool used = pw2_jac(M, S, W, i, jac, lj);
    for (int c = 0; c < 4; ++c) {
      J.pw_own[c * T + i] = mb[0] * jac[0][c] + mb[1] * jac[1][c] + mb[2] * jac[2][c];
      J.pw1_jac[c * T + i] = lj[c];
    }
    J.pw_w23bar[i] = used ? mb[0] * jac[0][4] + mb[1] * jac[1][4] + mb[2] * jac[2][4] : 0.;
  }
}

// The node maxima's adjoints, one owner per node (the first cell of pass
// 2's list around it, in the node-to-cell map's order, whose lowest vertex
// it is and whose node-maximum adjoint is nonzero): summed over those cells
// and routed to the cell attaining the maximum (its corner slot, read and
// cleared by k_adj_recon) or to the node's own height.  A cell is in the
// list when it is partially wet with its level not above every vertex.
__global__ void __launch_bounds__(256) k_adj_pwnode(DevMesh M, DevState S, DevWork W, DevAdj J) {
  const unsigned npw = *W.pw_count;
  const int64_t T = M.T;
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < npw; q += gridDim.x * blockDim.x) {
    const int i = W.pw_list[q];
    if (J.pw_w23bar[i] == 0.) continue;
    double zmx;
    const int n = lowest_node(M, i, zmx);
    int owner = -1;
    double sum = 0.;
    for (int r = M.n2c_off[n]; r < M.n2c_off[n + 1]; ++r) {
      const int c = M.n2c[r];
      if (W.cls[c] != 1) continue;
      if (lowest_node(M, c, zmx) != n || S.w[c] > zmx) continue;
      const double g = J.pw_w23bar[c];
      if (g == 0.) continue;
      if (owner < 0) owner = c;
      sum += g;
    }
    if (owner != i) continue;
    int arg;
    node_max_arg(M, W, n, node(M, n), arg);
    if (arg >= 0) {
      const int k = M.tp[arg] == n ? 0 : (M.tp[T + arg] == n ? 1 : 2);
      J.p1bar[k * T + arg] = sum;
    } else {
      J.zn_acc[n] += sum;
    }
  }
}

// ------------------------------------------------------- reconstruction --

// Reverse of k_recon1 / k_recon2 for cell i: gathers the MUSCL adjoint of
// the cell (its own update and the slots of its three edges), then maps it
// to the w adjoints of the cell and its neighbours and to corner z.
template <bool NT>
__global__ void __launch_bounds__(256) k_adj_recon(DevMesh M, DevState S, DevWork W, DevAdj J, double dts) {
  const int64_t i64 = global_index();  // XCD-aware, as the forward kernels
  if (i64 >= M.T) return;
  const int i = (int)i64;
  const int64_t T = M.T, Ecnt = M.E;
  const CVtx cvi = M.cv[i];
  const uint32_t fl = cvi.fl;
  const int cls = W.cls[i];
  // the update's gradient adjoint (its level and velocity parts it applied
  // itself); partially wet cells go through k_adj_pw's result
  double mb[5] = {0., 0., 0., 0., 0.};
  if (cls != 1) {
    mb[1] = J.m_cell[T + i];
    mb[2] = J.m_cell[2 * T + i];
  }
  double zc[3] = {0., 0., 0.};
  double wb = 0., bcb = 0.;
  // every load of the edge part is issued before any branch on loaded data
  // and before any store: the draining depth term, the edges' tags and both
  // sides' w adjoints and this side's velocity adjoints
  const double dtdi = W.dtd[i], dhi = J.drain_h[i];
  int ek[3], tag[3];
  double s0[3], s1[3], su[3], sv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int e = M.te[k * T + i];
    const int side = sgn_pos(fl, k) ? 0 : 1;
    ek[k] = e;
    tag[k] = M.et1[e];
    s0[k] = J.sl_edge[e];
    s1[k] = J.sl_edge[Ecnt + e];
    su[k] = J.sl_edge[(1 * 2 + side) * Ecnt + e];
    sv[k] = J.sl_edge[(2 * 2 + side) * Ecnt + e];
  }
  if (dtdi < dts) {  // the depth term of the draining dt (k_adj_flux)
    wb += dhi;
    bcb -= dhi;
  }
  const V3 p0 = node(M, cvi.v0), p1 = node(M, cvi.v1), p2 = node(M, cvi.v2);
  const CXY ci = M.cxy[i];
  const double cxi = ci.x, cyi = ci.y;
  {
    // the slot values of the cell's three edges (AtPoint at the midpoints)
    const Muscl mu = load_muscl(S, W, i);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const bool own = sgn_pos(fl, k);
      if (tag[k] == kSolidWall) {  // the edge's tag (tt is -1 on every boundary edge)
        const double hf = s0[k];  // side 0 w: the depth adjoint
        wb += hf;
        bcb -= hf;
        continue;
      }
      // the flux's own z term, once: -(w adjoints of both sides)
      double ez = own ? -(s0[k] + s1[k]) : 0.;
      slot_bar(mu, cxi, cyi, edge_mid(p0, p1, p2, fl, k), own ? s0[k] : s1[k], su[k], sv[k], mb, ez);
      zc[ep0_local(fl, k)] += 0.5 * ez;
      zc[ep1_local(fl, k)] += 0.5 * ez;
    }
  }
  const double w = S.w[i];
  double wnb[3] = {0., 0., 0.};
  // the node maxima this cell attains (k_adj_pwnode): adjoint of its pass-1
  // reconstruction's value at those vertices
  double mp[3] = {0., 0., 0.};
  bool p1clear[3] = {false, false, false};  // consumed node-maximum adjoints, cleared at the end
  if (*W.pw_count) {
    const V3 pv[3] = {p0, p1, p2};
    for (int k = 0; k < 3; ++k) {
      const double pk = J.p1bar[k * T + i];
      if (pk != 0.) {
        p1clear[k] = true;
        mp[0] += pk;
        mp[1] += pk * (pv[k].x - cxi);
        mp[2] += pk * (pv[k].y - cyi);
      }
    }
  }
  if (cls != 1)  // pass 1 is the final reconstruction
    for (int q = 0; q < 3; ++q) mb[q] += mp[q];
  if (cls == 0) {
    // ReconstructDryCell: level b_cell, gradient the bathymetry slope
    bcb += mb[0];
    for (int k = 0; k < 3; ++k) {
      V3 q0 = p0, q1 = p1, q2 = p2;
      q0.z = k == 0 ? 1. : 0.;
      q1.z = k == 1 ? 1. : 0.;
      q2.z = k == 2 ? 1. : 0.;
      double gx, gy;
      gradient(q0, q1, q2, gx, gy);
      zc[k] += mb[1] * gx + mb[2] * gy;
    }
  } else {
    const double zmax = dmax(dmax(p0.z, p1.z), p2.z);
    if (cls == 1 && w > zmax) {
      // not in pass 2's list: the level is w with a zero gradient, also as
      // the pass-1 value of the node maxima this cell attains
      wb += mb[0] + mp[0];
    } else if (cls == 1) {
      // PartWet2 / PartWet1 (k_adj_pw), and the PartWet1 level as the
      // pass-1 value of the node maxima this cell attains
      wb += J.pw_own[i];
      for (int k = 0; k < 3; ++k) zc[k] += J.pw_own[(1 + k) * T + i];
      if (mp[0] != 0.) {
        wb += mp[0] * J.pw1_jac[i];
        for (int k = 0; k < 3; ++k) zc[k] += mp[0] * J.pw1_jac[(1 + k) * T + i];
      }
    } else {
      // ReconstructFullWetCell (src/MUSCLObject.cpp:38-84): w0 = w; the
      // gradient is the plane through the neighbour values unless a
      // neighbour is dry or the positivity / TVD limiter zeroed it.  The
      // decisions are recomputed as k_recon1 takes them.
      wb += mb[0];
      V3 pts[3], E[3];
      int kind[3];
      bool any_dry = false;
      for (int k = 0; k < 3; ++k) {
        const int j = M.tt[k * T + i];  // full-wet cells are interior
        const NRec rj = M.nrec[j];  // highest vertex +inf on boundary cells
        const double wj = S.w[j];
        E[k] = edge_mid(p0, p1, p2, fl, k);
        kind[k] = 0;
        if (rj.zmx < wj) {
          pts[k] = {rj.cx, rj.cy, wj};
        } else if (!is_wet(wj - rj.bc)) {
          any_dry = true;
          pts[k] = {0., 0., 0.};
        } else {
          kind[k] = 1;
          const CellRec rf = M.crec[j];  // partially wet neighbour: the full record
          const double w0j = partwet1_w(rf.zmin, rf.zmin, rf.zmax, rf.bc, wj);
          const double dx = E[k].x - rj.cx, dy = E[k].y - rj.cy;
          double a = w0j + (0. * dx + 0. * dy);
          if (!((a - E[k].z) >= 0)) {
            a = E[k].z;
            kind[k] = 2;  // clamped to the edge bathymetry
          } else if (!(wj >= rf.zmax)) {
            kind[k] = 3;  // cube-root / bisection level of the neighbour
          }
          pts[k] = {E[k].x, E[k].y, 0.5 * (w + a)};
        }
      }
      bool active = false;
      if (!any_dry) {
        double gx, gy;
        gradient(pts[0], pts[1], pts[2], gx, gy);
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
        bool tvd = true;
        for (int k = 0; k < 3; ++k) {
          const int j = M.tt[k * T + i];
          const double wj = S.w[j];
          double vtmin = dmin(w, wj), vtmax = dmax(w, wj);
          double vek = w + (gx * (E[k].x - cxi) + gy * (E[k].y - cyi));
          if (!((vtmin <= vek) && (vek <= vtmax))) tvd = false;
        }
        active = all_wet && tvd;
      }
      if (active) {
        for (int k = 0; k < 3; ++k) {
          V3 q0 = pts[0], q1 = pts[1], q2 = pts[2];
          q0.z = k == 0 ? 1. : 0.;
          q1.z = k == 1 ? 1. : 0.;
          q2.z = k == 2 ? 1. : 0.;
          double gx, gy;
          gradient(q0, q1, q2, gx, gy);
          const double vb = mb[1] * gx + mb[2] * gy;  // adjoint of the k-th plane value
          if (kind[k] == 0) {
            wnb[k] = vb;
          } else {
            wb += 0.5 * vb;
            if (kind[k] == 1) wnb[k] = 0.5 * vb;  // the neighbour's level is its w
            else if (kind[k] == 2) {
              zc[ep0_local(fl, k)] += 0.25 * vb;  // a = E.z
              zc[ep1_local(fl, k)] += 0.25 * vb;
            } else {
              // the neighbour's PartWet1 level: its w (gathered by it), the
              // two vertices it shares with this cell (own corners) and the
              // opposite one (per edge side, folded in k_adj_z)
              const int j = M.tt[k * T + i];  // partially wet: in pass 2's list
              const double lb = 0.5 * vb;
              wnb[k] = lb * J.pw1_jac[j];
              const int e = M.te[k * T + i];
              const int sj = M.et0[e] == j ? 0 : 1;
              for (int m = 0; m < 3; ++m) {
                const int nj = M.tp[m * T + j];
                const int mi = M.tp[i] == nj ? 0 : (M.tp[T + i] == nj ? 1 : (M.tp[2 * T + i] == nj ? 2 : -1));
                const double g = lb * J.pw1_jac[(1 + m) * T + j];
                if (mi >= 0) zc[mi] += g;
                else J.zopp_acc[sj * Ecnt + e] += g;
              }
            }
          }
        }
      }
    }
  }
  // the stores, after every load (the compiler may not move a load of one
  // buffer past a store to another)
  if (cls != 0) {
    J.sbar[T + i] += mb[3];
    J.sbar[2 * T + i] += mb[4];
  }
  for (int k = 0; k < 3; ++k)
    if (p1clear[k]) J.p1bar[k * T + i] = 0.;
  for (int k = 0; k < 3; ++k) out_store<NT>(&J.w_nb[k * T + i], wnb[k]);
  out_store<NT>(&J.sbar[i], J.sbar[i] + wb);
  const double c = 1. / 3.;
  for (int k = 0; k < 3; ++k) out_store<NT>(&J.zc[k * T + i], J.zc[k * T + i] + (zc[k] + c * bcb));
}

// the neighbours' contributions to the w adjoint of cell c (after the last
// reverse stage; earlier ones are gathered by the next k_adj_update)
__global__ void __launch_bounds__(256) k_adj_gather(DevMesh M, DevAdj J) {
  const int64_t c64 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c64 >= M.T) return;
  const int c = (int)c64;
  J.sbar[c] += gather_nb(M, J.back, J.w_nb, c);
}

// node bathymetry adjoint: the corners of every incident cell, in the
// node -> cell map's order
__global__ void __launch_bounds__(256) k_adj_z(DevMesh M, DevAdj J, double* zbar) {
  const int64_t n64 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n64 >= M.V) return;
  const int n = (int)n64;
  const int64_t T = M.T;
  double s = 0.;
  for (int q = M.n2c_off[n]; q < M.n2c_off[n + 1]; ++q) {
    const int c = M.n2c[q];
    const int k = M.tp[c] == n ? 0 : (M.tp[T + c] == n ? 1 : 2);
    s += J.zc[k * T + c];
    // the side of c on its edge opposite this vertex
    const uint32_t fl = M.cflag[c];
    for (int kk = 0; kk < 3; ++kk)
      if (ep0_local(fl, kk) != k && ep1_local(fl, kk) != k)
        s += J.zopp_acc[(sgn_pos(fl, kk) ? 0 : 1) * M.E + M.te[kk * T + c]];
  }
  s += J.zn_acc[n];
  zbar[n] += s;
}

// the back-index and edge-local-index tables of the reverse kernels
__global__ void __launch_bounds__(256) k_adj_tables(DevMesh M, DevAdj J) {
  const int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t T = M.T;
  if (x < T) {
    const int c = (int)x;
    for (int k = 0; k < 3; ++k) {
      const int j = M.tt[k * T + c];
      J.back[k * T + c] = (int8_t)(j >= 0 ? back_index(M, j, c) : 0);
    }
  }
  if (x < M.E) {
    const int e = (int)x;
    const int t = M.et1[e];
    J.eloc[e] = (int8_t)local_edge(M, M.et0[e], e);
    J.eloc[M.E + e] = (int8_t)(t >= 0 ? local_edge(M, t, e) : 0);
  }
}

static inline dim3 grid_plain(int64_t n) {
  int64_t nb = (n + kBlock - 1) / kBlock;
  return dim3((unsigned)(nb > 0 ? nb : 1));
}

hipError_t launch_adjoint_stage(const DevMesh& M, const DevState& S, const DevState& U0, const DevWork& W,
                                const StageArgs& A, const DevAdj& J, int flux, int ws, int opts,
                                const double* wnb_in, int u0mode, unsigned int npw, hipStream_t st) {
  const double dts = A.coef * A.dt_fixed;
  if (M.nt) hipLaunchKernelGGL(k_adj_update<true>, grid_for(M.T), dim3(256), 0, st, M, S, U0, W, A, J, wnb_in, u0mode);
  else hipLaunchKernelGGL(k_adj_update<false>, grid_for(M.T), dim3(256), 0, st, M, S, U0, W, A, J, wnb_in, u0mode);
#define SWE_ADJ_FLUX(F, Wv, L)                                                                   \
  if (flux == F && ws == Wv && len == L)                                                         \
    { if (M.nt) hipLaunchKernelGGL((k_adj_flux<F, Wv, L, true>), grid_for(M.E), dim3(256), 0, st, M, S, W, J, dts); \
      else hipLaunchKernelGGL((k_adj_flux<F, Wv, L, false>), grid_for(M.E), dim3(256), 0, st, M, S, W, J, dts); }
  for (int li = 0; li < 2; ++li) {
    const bool len = li == 1;
    if (len != ((opts & kOptDrainLength) != 0)) continue;
    SWE_ADJ_FLUX(0, 0, false) SWE_ADJ_FLUX(0, 1, false) SWE_ADJ_FLUX(0, 2, false)
    SWE_ADJ_FLUX(1, 0, false) SWE_ADJ_FLUX(1, 1, false) SWE_ADJ_FLUX(1, 2, false)
    SWE_ADJ_FLUX(0, 0, true) SWE_ADJ_FLUX(0, 1, true) SWE_ADJ_FLUX(0, 2, true)
    SWE_ADJ_FLUX(1, 0, true) SWE_ADJ_FLUX(1, 1, true) SWE_ADJ_FLUX(1, 2, true)
  }
#undef SWE_ADJ_FLUX
  if (npw != 0u) {
    // partially wet cells: grid-stride over pass 2's list
    dim3 g = grid_plain(npw == kPwUnknown ? M.T : (int64_t)npw);
    if (g.x > 1024) g.x = 1024;
    hipLaunchKernelGGL(k_adj_pw, g, dim3(256), 0, st, M, S, W, J);
    hipLaunchKernelGGL(k_adj_pwnode, g, dim3(256), 0, st, M, S, W, J);
  }
  if (M.nt) hipLaunchKernelGGL(k_adj_recon<true>, grid_for(M.T), dim3(256), 0, st, M, S, W, J, dts);
  else hipLaunchKernelGGL(k_adj_recon<false>, grid_for(M.T), dim3(256), 0, st, M, S, W, J, dts);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_slot_counts(unsigned int* first, int64_t stride, int n,
                                                     unsigned int* out) {
  const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (x >= n) return;
  unsigned int* c = (unsigned int*)((char*)first + (int64_t)x * stride);
  if (out) out[x] = *c;
  else *c = 0u;
}

hipError_t launch_slot_counts(unsigned int* first, int64_t stride, int n, unsigned int* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_slot_counts, grid_plain(n), dim3(256), 0, st, first, stride, n, out);
  return hipGetLastError();
}

hipError_t launch_adjoint_tables(const DevMesh& M, const DevAdj& J, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_tables, grid_plain(M.T > M.E ? M.T : M.E), dim3(256), 0, st, M, J);
  return hipGetLastError();
}

hipError_t launch_adjoint_finish(const DevMesh& M, const DevAdj& J, double* zbar, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_gather, grid_plain(M.T), dim3(256), 0, st, M, J);
  if (zbar) hipLaunchKernelGGL(k_adj_z, grid_plain(M.V), dim3(256), 0, st, M, J, zbar);
  return hipGetLastError();
}

}  // namespace swe
