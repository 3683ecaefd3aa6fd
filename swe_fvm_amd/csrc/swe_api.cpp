// Solver runtime: device-resident mesh and state, stage sequencing of the
// SSP Runge-Kutta step on one HIP stream, getters serving the reference's
// SpaceDisc accessors, and event-based timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/swe_fvm.h"
#include "swe_internal.h"
#include "swe_launch.h"
#include "swe_tiles.h"

using namespace swe;

struct swe_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  int flux = 1, ws = 2, opts = 0;
  int flux_items = 2, upd_items = 1;  // work items per thread (swe_autotune)
  int flux_block = 256, upd_block = 256;  // threads per workgroup (swe_autotune)
  double cor = 0., tau = 0.;
  int64_t T = 0, E = 0, V = 0;
  std::vector<void*> allocs;
  DevMesh M{};
  DevGeomOut G{};  // writable static geometry (swe_set_bathymetry)
  // three state copies, each one allocation of 3 * Tpad doubles (Tpad = T
  // rounded up to kLayoutMaxBlock cells) holding w, u, v in the layout
  // lay_b (0: SoA, else blocks of lay_b cells); sw, su, sv point into it
  double* sbuf[3] = {nullptr, nullptr, nullptr};
  int64_t Tpad = 0;
  int lay_b = 0;
  double* sw[3] = {nullptr, nullptr, nullptr};
  double* su[3] = {nullptr, nullptr, nullptr};
  double* sv[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  DevWork W{};
  unsigned long long* rbits = nullptr;  // [4]
  int last_r_slot = 3;
  double* dts = nullptr;
  int64_t dts_cap = 0;
  int64_t last_nsteps = 0;
  // the last swe_adjoint_run: forward sweep, the reverse sweep's
  // recomputation, the rest of the reverse sweep (milliseconds)
  double adj_ms[3] = {0., 0., 0.};
  // batched ensemble (n_inst > 0)
  int n_inst = 0;
  DevBatch B{};
  std::vector<int64_t> cell_off, edge_off, node_off;
  hipGraphExec_t graph = nullptr;
  int graph_scheme = -1;
  double graph_cfl = 0., graph_tend = 0.;
  int64_t graph_c0 = -1, graph_c1 = -1;
  int graph_cur = -1;  // state buffer the captured launches start from
  double* scratch = nullptr;  // this device's static scratch array (swe_integrals_noalloc)
  // fused flux / draining / update kernel over tiles (single meshes whose
  // tiles fit its LDS): used for every stage whose step size is known
  DevTiles tiles{};
  bool fused_ok = false, fused_on = true;
  // after a fused stage W holds no fluxes: they are recomputed from the
  // stage's input buffer on demand (materialize_fluxes)
  bool flux_stale = false;
  int stale_in = -1, stale_slot = 0;
};

namespace {

// the reverse sweep recomputes a stage with the storing fused kernel below
// this size, with the separate flux kernel and then the tile kernel reading
// its fluxes (storing the draining steps) above
constexpr int64_t kFusedRecomputeMaxCells = int64_t(1) << 20;

// entry points of the single-mesh step (one CFL slot per stage) refuse
// ensemble handles, whose flux pass keeps one slot per member
int single_mesh(const swe_solver* s) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (s->n_inst > 0) return fail(SWE_ERR_ARG, "ensemble solver: use swe_batch_run");
  return SWE_OK;
}

// the linear friction coefficient of the update (0 unless SWE_OPT_FRICTION)
double friction(const swe_solver* s) { return (s->opts & SWE_OPT_FRICTION) ? s->tau : 0.; }

int dev_fail(hipError_t e, const char* what) {
  return fail(SWE_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
hipError_t dalloc(swe_solver* s, T** p, int64_t n) {
  size_t bytes = sizeof(T) * size_t(std::max<int64_t>(n, 1));
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e == hipSuccess) s->allocs.push_back(*p);
  return e;
}

uint32_t layout_mask(int block) { return block == 0 ? kLayoutSoA : uint32_t(block - 1); }
// offsets of the u and v planes from w in a state allocation
int64_t layout_plane(const swe_solver* s) { return s->lay_b == 0 ? s->Tpad : int64_t(s->lay_b); }
void layout_ptrs(swe_solver* s) {
  const int64_t d = layout_plane(s);
  for (int b = 0; b < 3; ++b) {
    s->sw[b] = s->sbuf[b];
    s->su[b] = s->sbuf[b] ? s->sbuf[b] + d : nullptr;
    s->sv[b] = s->sbuf[b] ? s->sbuf[b] + 2 * d : nullptr;
  }
}
DevState state_of(const swe_solver* s, int b) {
  return DevState{s->sw[b], s->su[b], s->sv[b], layout_mask(s->lay_b)};
}
DevOut out_of(const swe_solver* s, int b) { return DevOut{s->sw[b], s->su[b], s->sv[b], layout_mask(s->lay_b)}; }

#define SWE_CHECK(call, what)                 \
  do {                                        \
    hipError_t e_ = (call);                   \
    if (e_ != hipSuccess) return dev_fail(e_, what); \
  } while (0)

int set_device(swe_solver* s) {
  SWE_CHECK(hipSetDevice(s->device), "hipSetDevice");
  return SWE_OK;
}

// The fluxes and CFL minimum of the last fused stage, recomputed into W and
// its r slot from that stage's input (still intact: no stage writes its own
// input) and its reconstruction, for the getters that read them and before
// W's reconstruction is overwritten.  Bit-identical to what the separate
// flux kernel would have left.
int materialize_fluxes(swe_solver* s) {
  if (!s->flux_stale) return SWE_OK;
  SWE_CHECK(launch_flux(s->M, state_of(s, s->stale_in), s->W, s->flux, s->ws, s->cor, s->rbits + s->stale_slot,
                        s->flux_items, s->flux_block, s->stream), "flux");
  s->flux_stale = false;
  return SWE_OK;
}

// a stage takes the fused kernel when its step size is known before the
// flux pass: a fixed dt, or a later stage of a CFL step (slot > 0)
bool use_fused(const swe_solver* s, double dt_fixed, int slot) {
  return s->fused_ok && s->fused_on && (dt_fixed > 0. || slot > 0);
}
// otherwise (the first stage of a CFL step) the separate flux pass finds the
// step size, and the draining and update phases still run as one tile
// kernel reading its fluxes
bool use_fused_after_flux(const swe_solver* s) { return s->fused_ok && s->fused_on; }

// One stage: ComputeInterfaceValues, ComputeFluxes, update loop.
int run_stage(swe_solver* s, int in, int u0, int out, double a0, double a1, double coef,
              double dt_fixed, double cfl, int slot, double* dts_log) {
  DevState S = state_of(s, in);
  SWE_CHECK(launch_recon1(s->M, S, s->W, s->opts, s->stream), "recon1");
  SWE_CHECK(launch_recon2(s->M, S, s->W, s->opts, s->stream), "recon2");
  StageArgs A{a0, a1, coef, dt_fixed, cfl, s->cor, s->rbits, dts_log, friction(s)};
  if (use_fused(s, dt_fixed, slot)) {
    SWE_CHECK(launch_fused(s->M, s->tiles, S, state_of(s, u0), out_of(s, out), s->W, A, s->flux, s->ws, s->opts,
                           s->stream), "fused");
    s->flux_stale = true;
    s->stale_in = in;
    s->stale_slot = slot;
  } else {
    SWE_CHECK(launch_flux(s->M, S, s->W, s->flux, s->ws, s->cor, s->rbits + slot, s->flux_items, s->flux_block, s->stream), "flux");
    s->flux_stale = false;
    if (use_fused_after_flux(s)) {
      SWE_CHECK(launch_fused(s->M, s->tiles, S, state_of(s, u0), out_of(s, out), s->W, A, s->flux, s->ws, s->opts,
                             s->stream, false, true), "fused from fluxes");
    } else {
      SWE_CHECK(launch_drain(s->M, S, s->W, s->opts, s->stream), "drain");
      SWE_CHECK(launch_update(s->M, S, state_of(s, u0), out_of(s, out), s->W, A, s->upd_items, s->upd_block, s->stream), "update");
    }
  }
  s->last_r_slot = slot;
  return SWE_OK;
}

int run_step(swe_solver* s, int scheme, double dt, double cfl, double* dts_log) {
  int a = s->cur;
  int b = (a + 1) % 3, c = (a + 2) % 3;
  int rc;
  if (scheme == SWE_SCHEME_EULER) {
    rc = run_stage(s, a, a, b, 0., 0., 1., dt, cfl, 0, dts_log);
    if (rc) return rc;
    s->cur = b;
  } else if (scheme == SWE_SCHEME_SSPRK2) {
    rc = run_stage(s, a, a, b, 0., 0., 1., dt, cfl, 0, dts_log);
    if (rc) return rc;
    rc = run_stage(s, b, a, a, 0.5, 0.5, 0.5, dt, cfl, 1, nullptr);
    if (rc) return rc;
  } else {
    rc = run_stage(s, a, a, b, 0., 0., 1., dt, cfl, 0, dts_log);
    if (rc) return rc;
    rc = run_stage(s, b, a, c, 0.75, 0.25, 0.25, dt, cfl, 1, nullptr);
    if (rc) return rc;
    rc = run_stage(s, c, a, a, 1. / 3., 2. / 3., 2. / 3., dt, cfl, 2, nullptr);
    if (rc) return rc;
  }
  return SWE_OK;
}

}  // namespace

static int create_impl(const Topo& m, const double* xyz, int flux, int wavespeed, double cor,
                       double tau, int options, int device, const std::vector<int>* cell_inst,
                       int n_inst, swe_solver** out) {
  if (!xyz || !out) return fail(SWE_ERR_ARG, "null argument");
  if (flux < 0 || flux > 1 || wavespeed < 0 || wavespeed > 2)
    return fail(SWE_ERR_ARG, "unknown flux or wave-speed estimate");
  if (options & ~(SWE_OPT_HEAD_GRADIENT | SWE_OPT_HEAD_PARTWET2 | SWE_OPT_FRICTION | SWE_OPT_DRAIN_LENGTH |
                  SWE_OPT_EXACT_BALANCE))
    return fail(SWE_ERR_ARG, "unknown option bits");
  if (m.T > 0x7fffffff || m.E > 0x7fffffff || m.V > 0x7fffffff)
    return fail(SWE_ERR_ARG, "mesh too large for int32 device indices");
  if (m.T <= 0) return fail(SWE_ERR_ARG, "mesh without triangles");
  const int64_t T = m.T, E = m.E, V = m.V;
  // host-side validation and SoA packing
  std::vector<int> tp(size_t(3 * T)), te(size_t(3 * T)), tt(size_t(3 * T));
  std::vector<uint32_t> flag(size_t(T), 0u);
  std::vector<int> ep0(static_cast<size_t>(E)), ep1(static_cast<size_t>(E)), et0(static_cast<size_t>(E)), et1(static_cast<size_t>(E));
  for (int64_t e = 0; e < E; ++e) {
    int64_t a = m.ep[2 * e], b = m.ep[2 * e + 1], f = m.et[2 * e], t = m.et[2 * e + 1];
    if (a < 0 || a >= V || b < 0 || b >= V) return fail(SWE_ERR_ARG, "edge node out of range");
    if (f < 0 || f >= T || t >= T) return fail(SWE_ERR_ARG, "edge element out of range");
    if (t < -3)
      return fail(SWE_ERR_DOMAIN, "boundary tag other than SOLID_WALL (-1), FREE_FLOW (-2) or PERIODIC (-3)");
    if (t == -3) {
      const int64_t q = m.pe.empty() ? -1 : m.pe[size_t(e)];
      if (q < 0 || q >= E || m.et[2 * q + 1] != -3 || m.pe[size_t(q)] != e)
        return fail(SWE_ERR_DOMAIN, "PERIODIC edge without a partner (swe_topology_set_periodic)");
    }
    ep0[size_t(e)] = int(a); ep1[size_t(e)] = int(b); et0[size_t(e)] = int(f); et1[size_t(e)] = int(t);
  }
  std::vector<int> deg(size_t(V) + 1, 0);
  for (int64_t i = 0; i < T; ++i) {
    uint32_t fl = 0;
    for (int k = 0; k < 3; ++k) {
      int64_t p = m.tp[3 * i + k], e = m.te[3 * i + k], n = m.tt[3 * i + k];
      if (p < 0 || p >= V) return fail(SWE_ERR_ARG, "element node out of range");
      if (e < 0 || e >= E) return fail(SWE_ERR_ARG, "element edge out of range");
      if (n < -1 || n >= T) return fail(SWE_ERR_ARG, "element neighbour out of range");
      tp[size_t(k * T + i)] = int(p);
      te[size_t(k * T + i)] = int(e);
      tt[size_t(k * T + i)] = int(n);
      deg[size_t(p)]++;
    }
    for (int k = 0; k < 3; ++k) {
      int64_t e = m.te[3 * i + k];
      // the kernels index neighbours through tt: it must name the other cell
      // of edge te[k] (or -1 on a wall), as the reference builder writes it
      int64_t a = m.et[2 * e], b = m.et[2 * e + 1];
      if (a != i && b != i) return fail(SWE_ERR_ARG, "element edge does not list the element");
      int64_t other = (a == i) ? b : a;
      if (m.tt[3 * i + k] != (other < 0 ? -1 : other))
        return fail(SWE_ERR_ARG, "element neighbour disagrees with the edge's elements");
      if (m.et[2 * e] == i) fl |= 1u << k;
      if (m.et[2 * e + 1] < 0) fl |= 1u << 3;
      int la = -1, lb = -1;
      for (int q = 0; q < 3; ++q) {
        if (m.tp[3 * i + q] == m.ep[2 * e]) la = q;
        if (m.tp[3 * i + q] == m.ep[2 * e + 1]) lb = q;
      }
      if (la < 0 || lb < 0 || la == lb)
        return fail(SWE_ERR_ARG, "element edge does not join two of the element's nodes");
      fl |= uint32_t(la) << (4 + 4 * k);
      fl |= uint32_t(lb) << (6 + 4 * k);
    }
    flag[size_t(i)] = fl;
  }
  std::vector<int> n2c_off(size_t(V) + 1, 0), n2c(size_t(3 * T));
  for (int64_t n = 0; n < V; ++n) n2c_off[size_t(n) + 1] = n2c_off[size_t(n)] + deg[size_t(n)];
  {
    std::vector<int> fillp(n2c_off.begin(), n2c_off.end() - 1);
    for (int64_t i = 0; i < T; ++i)
      for (int k = 0; k < 3; ++k) n2c[size_t(fillp[size_t(m.tp[3 * i + k])]++)] = int(i);
  }
  std::vector<int> pe;  // partner edges (periodic meshes only)
  for (int64_t e = 0; e < E; ++e)
    if (et1[size_t(e)] == -3) {
      if (pe.empty()) pe.assign(size_t(E), -1);
      pe[size_t(e)] = int(m.pe[size_t(e)]);
    }
  std::vector<double> px(static_cast<size_t>(V)), py(static_cast<size_t>(V)), pz(static_cast<size_t>(V));
  for (int64_t n = 0; n < V; ++n) {
    px[size_t(n)] = xyz[3 * n]; py[size_t(n)] = xyz[3 * n + 1]; pz[size_t(n)] = xyz[3 * n + 2];
  }

  // tables are validated first, so mesh errors are reported without a device
  int ndev = 0;
  hipError_t he = hipGetDeviceCount(&ndev);
  if (he != hipSuccess || ndev <= 0) return fail(SWE_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(SWE_ERR_ARG, "invalid device ordinal");

  auto* s = new swe_solver;
  s->device = device;
  s->flux = flux; s->ws = wavespeed; s->opts = options; s->cor = cor; s->tau = tau;
  s->T = T; s->E = E; s->V = V;
  auto bail = [&](hipError_t e, const char* what) {
    int rc = dev_fail(e, what);
    swe_destroy(s);
    return rc;
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
  if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail(e, "hipStreamCreate");
  if ((e = scratch_address(&s->scratch)) != hipSuccess) return bail(e, "hipGetSymbolAddress(scratch)");
  double *dpx, *dpy, *dpz;
  int *dtp, *dte, *dtt, *dep0, *dep1, *det0, *det1, *doff, *dn2c;
  uint32_t* dfl;
#define ALLOC(ptr, n) if ((e = dalloc(s, &ptr, n)) != hipSuccess) return bail(e, "hipMalloc")
  ALLOC(dpx, V); ALLOC(dpy, V); ALLOC(dpz, V);
  ALLOC(dtp, 3 * T); ALLOC(dte, 3 * T); ALLOC(dtt, 3 * T); ALLOC(dfl, T);
  ALLOC(dep0, E); ALLOC(dep1, E); ALLOC(det0, E); ALLOC(det1, E);
  ALLOC(doff, V + 1); ALLOC(dn2c, 3 * T);
  s->Tpad = (T + kLayoutMaxBlock - 1) / kLayoutMaxBlock * kLayoutMaxBlock;
  for (int b = 0; b < 3; ++b) ALLOC(s->sbuf[b], 3 * s->Tpad);
  layout_ptrs(s);
  ALLOC(s->W.mw, T); ALLOC(s->W.mgx, T); ALLOC(s->W.mgy, T);
  ALLOC(s->W.wrec1, T);
  ALLOC(s->W.cls, T);
  ALLOC(s->W.f0, E); ALLOC(s->W.f1, E); ALLOC(s->W.f2, E);
  ALLOC(s->W.dtd, T);
  ALLOC(s->W.pw_list, T); ALLOC(s->W.pw_count, 1);
  ALLOC(s->rbits, 4);
  DevGeomOut G{};
  ALLOC(G.cx, T); ALLOC(G.cy, T); ALLOC(G.bc, T);
  ALLOC(G.area, T);
  ALLOC(G.edmin, E);
  ALLOC(G.crec, T);
  ALLOC(G.nrec, T);
  ALLOC(G.cxy, T);
  ALLOC(G.pn, V);
  ALLOC(G.cv, T);
#undef ALLOC
#define H2D(dst, src, n)                                                                     \
  if ((e = hipMemcpy(dst, src, sizeof(*(src)) * size_t(n), hipMemcpyHostToDevice)) != hipSuccess) \
    return bail(e, "hipMemcpy");
  H2D(dpx, px.data(), V); H2D(dpy, py.data(), V); H2D(dpz, pz.data(), V);
  H2D(dtp, tp.data(), 3 * T); H2D(dte, te.data(), 3 * T); H2D(dtt, tt.data(), 3 * T);
  H2D(dfl, flag.data(), T);
  H2D(dep0, ep0.data(), E); H2D(dep1, ep1.data(), E); H2D(det0, et0.data(), E); H2D(det1, et1.data(), E);
  H2D(doff, n2c_off.data(), V + 1); H2D(dn2c, n2c.data(), 3 * T);
#undef H2D
  unsigned long long one;
  double d1 = 1.0;
  std::memcpy(&one, &d1, 8);
  unsigned long long rinit[4] = {one, one, one, one};
  if ((e = hipMemcpy(s->rbits, rinit, sizeof(rinit), hipMemcpyHostToDevice)) != hipSuccess)
    return bail(e, "hipMemcpy");
  if ((e = hipMemset(s->W.pw_count, 0, sizeof(unsigned int))) != hipSuccess) return bail(e, "hipMemset");
  for (int b = 0; b < 3; ++b) {
    if ((e = hipMemset(s->sbuf[b], 0, sizeof(double) * size_t(3 * s->Tpad))) != hipSuccess) return bail(e, "hipMemset");
  }
  // the whole mesh is the visited window; cinst / bdt / bactive are set for
  // ensembles below
  s->M = DevMesh{dpx, dpy, dpz, dtp, dte, dtt, dfl, dep0, dep1, det0, det1, doff, dn2c, T, E, V,
                 G.cx, G.cy, G.bc, G.area, G.edmin, G.crec, nullptr, nullptr, nullptr, 0, T, 0, E,
                 T >= (int64_t(1) << 20) ? 1 : 0, nullptr, (options & SWE_OPT_EXACT_BALANCE) ? 1 : 0, G.nrec, G.cxy, G.pn, G.cv};
  s->G = G;
  if (!pe.empty()) {
    int* dpe;
    if ((e = dalloc(s, &dpe, E)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMemcpy(dpe, pe.data(), sizeof(int) * size_t(E), hipMemcpyHostToDevice)) != hipSuccess)
      return bail(e, "hipMemcpy");
    s->M.pe = dpe;
  }
  // the fused kernel's tiles hold no periodic partners: periodic meshes keep
  // the separate flux, draining and update kernels (ensembles tile their
  // stacked mesh; no edge joins two members)
  if (pe.empty()) {
    TileTables tt_host;
    if (build_tiles(T, tt.data(), te.data(), flag.data(), et0.data(), et1.data(), ep0.data(), ep1.data(), tt_host)) {
      uint32_t *dcode, *drcode;
      int *droff, *drcell, *devoff;
      EdgeRec* dev;
      int* devedge;
#define ALLOC(ptr, n) if ((e = dalloc(s, &ptr, n)) != hipSuccess) return bail(e, "hipMalloc")
#define H2D(dst, v)                                                                                      \
  if (!v.empty() && (e = hipMemcpy(dst, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice)) != hipSuccess) \
    return bail(e, "hipMemcpy");
      ALLOC(dcode, int64_t(tt_host.code.size())); ALLOC(droff, int64_t(tt_host.ring_off.size()));
      ALLOC(drcell, int64_t(tt_host.ring_cell.size())); ALLOC(drcode, int64_t(tt_host.ring_code.size()));
      ALLOC(devoff, int64_t(tt_host.ev_off.size())); ALLOC(dev, int64_t(tt_host.evrec.size()));
      ALLOC(devedge, int64_t(tt_host.evedge.size()));
      H2D(dcode, tt_host.code); H2D(droff, tt_host.ring_off); H2D(drcell, tt_host.ring_cell);
      H2D(drcode, tt_host.ring_code); H2D(devoff, tt_host.ev_off); H2D(dev, tt_host.evrec);
      H2D(devedge, tt_host.evedge);
#undef H2D
#undef ALLOC
      s->tiles = DevTiles{dcode, droff, drcell, drcode, devoff, dev, devedge, tt_host.ntiles};
      s->fused_ok = true;
    }
  }
  if ((e = launch_setup_geometry(s->M, G, s->stream)) != hipSuccess) return bail(e, "setup_geometry");
  if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
  if (n_inst > 0) {
    s->n_inst = n_inst;
    int* dci;
    DevBatch& B = s->B;
    B.n = n_inst;
#define ALLOC(ptr, n) if ((e = dalloc(s, &ptr, n)) != hipSuccess) return bail(e, "hipMalloc")
    ALLOC(dci, T);
    ALLOC(B.t, n_inst); ALLOC(B.dt, n_inst); ALLOC(B.active, n_inst); ALLOC(B.steps, n_inst);
    ALLOC(B.rbits, 4 * n_inst); ALLOC(B.n_active, 1);
#undef ALLOC
    if ((e = hipMemcpy(dci, cell_inst->data(), sizeof(int) * size_t(T), hipMemcpyHostToDevice)) != hipSuccess)
      return bail(e, "hipMemcpy");
    s->M.cinst = dci;
    s->M.bdt = B.dt;
    s->M.bactive = B.active;
    if (int rc = swe_batch_reset(s, 0.0)) { swe_destroy(s); return rc; }
  }
  *out = s;
  return SWE_OK;
}

extern "C" {

int swe_create(const swe_topology* topo, const double* xyz, int flux, int wavespeed, double cor,
               double tau, int options, int device, swe_solver** out) {
  if (!topo) return fail(SWE_ERR_ARG, "null topology");
  return create_impl(topo->m, xyz, flux, wavespeed, cor, tau, options, device, nullptr, 0, out);
}

int swe_batch_create(int64_t n, const swe_topology* const* topos, const double* const* xyzs,
                     int flux, int wavespeed, double cor, double tau, int options, int device,
                     swe_solver** out) {
  if (n <= 0 || n > (1 << 20) || !topos || !xyzs || !out) return fail(SWE_ERR_ARG, "invalid ensemble");
  Topo all;
  std::vector<double> xyz;
  std::vector<int> inst;
  std::vector<int64_t> coff(1, 0), eoff(1, 0), noff(1, 0);
  for (int64_t k = 0; k < n; ++k) {
    if (!topos[k] || !xyzs[k]) return fail(SWE_ERR_ARG, "null ensemble member");
    const Topo& m = topos[k]->m;
    if (m.T <= 0) return fail(SWE_ERR_ARG, "ensemble member without triangles");
    const int64_t V0 = all.V, E0 = all.E, T0 = all.T;
    for (int64_t q = 0; q < 2 * m.E; ++q) all.ep.push_back(m.ep[size_t(q)] + V0);
    for (int64_t q = 0; q < 2 * m.E; ++q) {
      int64_t c = m.et[size_t(q)];
      all.et.push_back(c >= 0 ? c + T0 : c);
    }
    for (int64_t q = 0; q < 3 * m.T; ++q) all.tp.push_back(m.tp[size_t(q)] + V0);
    for (int64_t q = 0; q < 3 * m.T; ++q) all.te.push_back(m.te[size_t(q)] + E0);
    for (int64_t q = 0; q < 3 * m.T; ++q) {
      int64_t c = m.tt[size_t(q)];
      all.tt.push_back(c >= 0 ? c + T0 : c);
    }
    if (!m.pe.empty() || !all.pe.empty()) {
      all.pe.resize(size_t(E0), -1);
      for (int64_t q = 0; q < m.E; ++q) {
        const int64_t c = m.pe.empty() ? -1 : m.pe[size_t(q)];
        all.pe.push_back(c >= 0 ? c + E0 : -1);
      }
    }
    xyz.insert(xyz.end(), xyzs[k], xyzs[k] + 3 * m.V);
    inst.insert(inst.end(), size_t(m.T), int(k));
    all.V += m.V; all.E += m.E; all.T += m.T;
    coff.push_back(all.T); eoff.push_back(all.E); noff.push_back(all.V);
  }
  int rc = create_impl(all, xyz.data(), flux, wavespeed, cor, tau, options, device, &inst, int(n), out);
  if (rc == SWE_OK) {
    (*out)->cell_off = coff;
    (*out)->edge_off = eoff;
    (*out)->node_off = noff;
  }
  return rc;
}

void swe_destroy(swe_solver* s) {  // teardown: failures cannot be reported
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->graph) (void)hipGraphExecDestroy(s->graph);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->dts) (void)hipFree(s->dts);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// the host image of a state allocation in the solver's layout (padding cells
// zero) from / to the caller's (T, 3) primitive array
int swe_set_state(swe_solver* s, const double* prim) {
  if (!s || !prim) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  std::vector<double> h(size_t(3 * s->Tpad), 0.);
  const uint32_t mask = layout_mask(s->lay_b);
  const int64_t d = layout_plane(s);
  for (int64_t i = 0; i < s->T; ++i) {
    const int64_t a = state_at(i, mask);
    h[size_t(a)] = prim[3 * i]; h[size_t(d + a)] = prim[3 * i + 1]; h[size_t(2 * d + a)] = prim[3 * i + 2];
  }
  SWE_CHECK(hipMemcpyAsync(s->sbuf[s->cur], h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s->stream),
            "hipMemcpy");
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  return SWE_OK;
}

int swe_get_state(swe_solver* s, double* prim) {
  if (!s || !prim) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  std::vector<double> h(size_t(3 * s->Tpad));
  SWE_CHECK(hipMemcpyAsync(h.data(), s->sbuf[s->cur], sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream),
            "hipMemcpy");
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  const uint32_t mask = layout_mask(s->lay_b);
  const int64_t d = layout_plane(s);
  for (int64_t i = 0; i < s->T; ++i) {
    const int64_t a = state_at(i, mask);
    prim[3 * i] = h[size_t(a)]; prim[3 * i + 1] = h[size_t(d + a)]; prim[3 * i + 2] = h[size_t(2 * d + a)];
  }
  return SWE_OK;
}

int swe_compute_interface_values(swe_solver* s) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  DevState S = state_of(s, s->cur);
  SWE_CHECK(launch_recon1(s->M, S, s->W, s->opts, s->stream), "recon1");
  SWE_CHECK(launch_recon2(s->M, S, s->W, s->opts, s->stream), "recon2");
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  return SWE_OK;
}

int swe_compute_fluxes(swe_solver* s, double* r) {
  if (int rc = single_mesh(s)) return rc;
  if (int rc = set_device(s)) return rc;
  DevState S = state_of(s, s->cur);
  SWE_CHECK(launch_flux(s->M, S, s->W, s->flux, s->ws, s->cor, s->rbits + 3, s->flux_items, s->flux_block, s->stream), "flux");
  s->flux_stale = false;
  s->last_r_slot = 3;
  unsigned long long bits;
  SWE_CHECK(hipMemcpyAsync(&bits, s->rbits + 3, 8, hipMemcpyDeviceToHost, s->stream), "hipMemcpy");
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  if (r) std::memcpy(r, &bits, 8);
  return SWE_OK;
}

static int download(swe_solver* s, void* dst, const void* src, size_t bytes) {
  SWE_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream), "hipMemcpy");
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  return SWE_OK;
}

static int edge_and_src(swe_solver* s, double* edg, double* src) {
  if (int rc = set_device(s)) return rc;
  double *de = nullptr, *ds = nullptr;
  size_t n = size_t(6 * std::max<int64_t>(s->E, 1));
  SWE_CHECK(hipMalloc((void**)&de, sizeof(double) * n), "hipMalloc");
  hipError_t e = hipMalloc((void**)&ds, sizeof(double) * n);
  if (e != hipSuccess) { (void)hipFree(de); return dev_fail(e, "hipMalloc"); }
  e = hipMemsetAsync(de, 0, sizeof(double) * n, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(ds, 0, sizeof(double) * n, s->stream);
  if (e == hipSuccess) e = launch_edge_field(s->M, state_of(s, s->cur), s->W, s->cor, de, ds, s->stream);
  int rc = SWE_OK;
  if (e != hipSuccess) rc = dev_fail(e, "edge_field");
  if (!rc && edg) rc = download(s, edg, de, sizeof(double) * size_t(6 * s->E));
  if (!rc && src) rc = download(s, src, ds, sizeof(double) * size_t(6 * s->E));
  (void)hipStreamSynchronize(s->stream);  // cleanup: errors were reported above
  (void)hipFree(de);
  (void)hipFree(ds);
  return rc;
}

int swe_get_edge_field(swe_solver* s, double* edg) {
  if (!s || !edg) return fail(SWE_ERR_ARG, "null argument");
  return edge_and_src(s, edg, nullptr);
}
int swe_get_src_field(swe_solver* s, double* src) {
  if (!s || !src) return fail(SWE_ERR_ARG, "null argument");
  return edge_and_src(s, nullptr, src);
}

int swe_get_fluxes(swe_solver* s, double* f) {
  if (!s || !f) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  std::vector<double> a(size_t(s->E)), b(size_t(s->E)), c(size_t(s->E));
  if (int rc = download(s, a.data(), s->W.f0, sizeof(double) * a.size())) return rc;
  if (int rc = download(s, b.data(), s->W.f1, sizeof(double) * b.size())) return rc;
  if (int rc = download(s, c.data(), s->W.f2, sizeof(double) * c.size())) return rc;
  for (int64_t e = 0; e < s->E; ++e) {
    f[3 * e] = a[size_t(e)]; f[3 * e + 1] = b[size_t(e)]; f[3 * e + 2] = c[size_t(e)];
  }
  return SWE_OK;
}

int swe_get_max_wp(swe_solver* s, double* max_wp) {
  if (!s || !max_wp) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  double* d = nullptr;
  SWE_CHECK(hipMalloc((void**)&d, sizeof(double) * size_t(std::max<int64_t>(s->V, 1))), "hipMalloc");
  hipError_t e = launch_max_wp(s->M, s->W, d, s->stream);
  int rc = (e != hipSuccess) ? dev_fail(e, "max_wp") : download(s, max_wp, d, sizeof(double) * size_t(s->V));
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  return rc;
}

int swe_get_cell_class(swe_solver* s, int8_t* cls) {
  if (!s || !cls) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  return download(s, cls, s->W.cls, size_t(s->T));
}

int swe_reconstruct(swe_solver* s, int64_t cell, int kind, double* out5) {
  if (int rc = single_mesh(s)) return rc;
  if (!out5) return fail(SWE_ERR_ARG, "null argument");
  if (cell < 0 || cell >= s->T) return fail(SWE_ERR_ARG, "cell index out of range");
  if (kind < 0 || kind > 3) return fail(SWE_ERR_ARG, "kind must be 0 (dry), 1, 2 (full wet) or 3");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  DevState S = state_of(s, s->cur);
  SWE_CHECK(launch_recon1(s->M, S, s->W, s->opts, s->stream), "recon1");
  SWE_CHECK(launch_recon2(s->M, S, s->W, s->opts, s->stream), "recon2");
  double* d = nullptr;
  SWE_CHECK(hipMalloc((void**)&d, 5 * sizeof(double)), "hipMalloc");
  hipError_t e = launch_reconstruct_one(s->M, S, s->W, int(cell), kind, s->opts, d, s->stream);
  int rc = (e != hipSuccess) ? dev_fail(e, "reconstruct") : download(s, out5, d, 5 * sizeof(double));
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  return rc;
}

int swe_edge_flux(swe_solver* s, int flux, int wavespeed, int64_t e, int64_t from, int64_t to, double* r,
                  double* out3) {
  if (int rc = single_mesh(s)) return rc;
  if (!out3) return fail(SWE_ERR_ARG, "null argument");
  if (flux < 0 || flux > 1 || wavespeed < 0 || wavespeed > 2)
    return fail(SWE_ERR_ARG, "unknown flux or wave-speed estimate");
  if (e < 0 || e >= s->E) return fail(SWE_ERR_ARG, "edge index out of range");
  if (from < 0 || from >= s->T || to >= s->T || to < -2 || to == from)
    return fail(SWE_ERR_ARG, "cell index out of range");
  if (int rc = set_device(s)) return rc;
  double* d = nullptr;
  SWE_CHECK(hipMalloc((void**)&d, 4 * sizeof(double)), "hipMalloc");
  hipError_t he = launch_edge_flux_one(s->M, state_of(s, s->cur), s->W, flux, wavespeed, int(e), int(from),
                                       int(to), s->cor, d, s->stream);
  double o[4] = {0., 0., 0., 0.};
  int rc = (he != hipSuccess) ? dev_fail(he, "edge_flux") : download(s, o, d, 4 * sizeof(double));
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (rc) return rc;
  out3[0] = o[0]; out3[1] = o[1]; out3[2] = o[2];
  if (r && o[3] < *r) *r = o[3];  // include/Fluxes.h:42-48: r = min(r, candidate)
  return SWE_OK;
}

int swe_rhs(swe_solver* s, int64_t cell, double dt, double* out3) {
  if (!s || !out3) return fail(SWE_ERR_ARG, "null argument");
  if (cell < 0 || cell >= s->T) return fail(SWE_ERR_ARG, "cell index out of range");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  DevState S = state_of(s, s->cur);
  double* d = nullptr;
  SWE_CHECK(hipMalloc((void**)&d, 3 * sizeof(double)), "hipMalloc");
  hipError_t e = launch_drain(s->M, S, s->W, s->opts, s->stream);
  if (e == hipSuccess) e = launch_rhs_one(s->M, S, s->W, int(cell), dt, s->cor, friction(s), d, s->stream);
  int rc = (e != hipSuccess) ? dev_fail(e, "rhs") : download(s, out3, d, 3 * sizeof(double));
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  return rc;
}

int swe_run(swe_solver* s, int scheme, double dt, double cfl, int64_t nsteps) {
  if (int rc = single_mesh(s)) return rc;
  if (scheme < 0 || scheme > 2) return fail(SWE_ERR_ARG, "unknown scheme");
  if (nsteps < 0) return fail(SWE_ERR_ARG, "negative step count");
  if (!(dt > 0.) && !(cfl > 0.)) return fail(SWE_ERR_ARG, "CFL mode needs cfl > 0");
  if (int rc = set_device(s)) return rc;
  if (nsteps > s->dts_cap) {
    if (s->dts) { (void)hipStreamSynchronize(s->stream); (void)hipFree(s->dts); s->dts = nullptr; }
    SWE_CHECK(hipMalloc((void**)&s->dts, sizeof(double) * size_t(nsteps)), "hipMalloc");
    s->dts_cap = nsteps;
  }
  for (int64_t k = 0; k < nsteps; ++k)
    if (int rc = run_step(s, scheme, dt, cfl, s->dts + k)) return rc;
  s->last_nsteps = nsteps;
  return SWE_OK;
}

int swe_get_dts(swe_solver* s, double* dts, int64_t nsteps) {
  if (!s || !dts) return fail(SWE_ERR_ARG, "null argument");
  if (nsteps > s->last_nsteps) return fail(SWE_ERR_ARG, "more steps requested than run");
  if (int rc = set_device(s)) return rc;
  return download(s, dts, s->dts, sizeof(double) * size_t(nsteps));
}

int swe_synchronize(swe_solver* s) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (int rc = set_device(s)) return rc;
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  return SWE_OK;
}

int swe_cfl_dt(swe_solver* s, double* dt) {
  if (int rc = single_mesh(s)) return rc;
  if (!dt) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  unsigned long long bits;
  if (int rc = download(s, &bits, s->rbits + s->last_r_slot, 8)) return rc;
  double r;
  std::memcpy(&r, &bits, 8);
  *dt = 0.15 * r;  // include/TimeDisc.h:13,22
  return SWE_OK;
}

int swe_min_len_to_wavespeed(swe_solver* s, double* r) {
  if (int rc = single_mesh(s)) return rc;
  if (!r) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  unsigned long long bits;
  if (int rc = download(s, &bits, s->rbits + s->last_r_slot, 8)) return rc;
  std::memcpy(r, &bits, 8);
  return SWE_OK;
}

int64_t swe_num_cells(const swe_solver* s) { return s ? s->T : -1; }
int64_t swe_num_edges(const swe_solver* s) { return s ? s->E : -1; }
int64_t swe_num_nodes(const swe_solver* s) { return s ? s->V : -1; }

int swe_device_state(swe_solver* s, double** w, double** u, double** v) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (s->lay_b != 0) return fail(SWE_ERR_ARG, "the state is not in the SoA layout (swe_set_layout(s, 0))");
  if (w) *w = s->sw[s->cur];
  if (u) *u = s->su[s->cur];
  if (v) *v = s->sv[s->cur];
  return SWE_OK;
}

static int batch_step(swe_solver* s, int scheme, double cfl, double t_end) {
  // Euler ping-pongs between two buffers so that a two-step graph returns
  // the state to the buffer it started from
  const int a = s->cur;
  const int b = (scheme == SWE_SCHEME_EULER) ? (a == 0 ? 1 : 0) : (a + 1) % 3;
  const int c = (a + 2) % 3;
  struct St { int in, u0, out; double a0, a1, coef; };
  St st[3];
  int nst;
  if (scheme == SWE_SCHEME_EULER) {
    st[0] = {a, a, b, 0., 0., 1.}; nst = 1;
  } else if (scheme == SWE_SCHEME_SSPRK2) {
    st[0] = {a, a, b, 0., 0., 1.}; st[1] = {b, a, a, .5, .5, .5}; nst = 2;
  } else {
    st[0] = {a, a, b, 0., 0., 1.}; st[1] = {b, a, c, .75, .25, .25};
    st[2] = {c, a, a, 1. / 3., 2. / 3., 2. / 3.}; nst = 3;
  }
  for (int q = 0; q < nst; ++q) {
    DevState S = state_of(s, st[q].in);
    StageArgs A{st[q].a0, st[q].a1, st[q].coef, 0., cfl, s->cor, nullptr, nullptr, friction(s)};
    if (q > 0 && use_fused(s, 0., q)) {  // the members' step sizes are known: one fused kernel
      SWE_CHECK(launch_recon1(s->M, S, s->W, s->opts, s->stream), "recon1");
      SWE_CHECK(launch_recon2(s->M, S, s->W, s->opts, s->stream), "recon2");
      SWE_CHECK(launch_fused(s->M, s->tiles, S, state_of(s, st[q].u0), out_of(s, st[q].out), s->W, A, s->flux, s->ws,
                             s->opts, s->stream), "fused");
      continue;
    }
    SWE_CHECK(launch_batch_reset_r(s->B, q, s->stream), "batch_reset_r");
    SWE_CHECK(launch_recon1(s->M, S, s->W, s->opts, s->stream), "recon1");
    SWE_CHECK(launch_recon2(s->M, S, s->W, s->opts, s->stream), "recon2");
    SWE_CHECK(launch_flux(s->M, S, s->W, s->flux, s->ws, s->cor, s->B.rbits + q, s->flux_items, s->flux_block, s->stream), "flux");
    if (q == 0) SWE_CHECK(launch_batch_dt(s->B, cfl, t_end, s->stream), "batch_dt");
    if (use_fused_after_flux(s)) {
      SWE_CHECK(launch_fused(s->M, s->tiles, S, state_of(s, st[q].u0), out_of(s, st[q].out), s->W, A, s->flux, s->ws,
                             s->opts, s->stream, false, true), "fused from fluxes");
      continue;
    }
    SWE_CHECK(launch_drain(s->M, S, s->W, s->opts, s->stream), "drain");
    SWE_CHECK(launch_update(s->M, S, state_of(s, st[q].u0), out_of(s, st[q].out), s->W, A, s->upd_items, s->upd_block, s->stream),
              "update");
  }
  SWE_CHECK(launch_batch_advance(s->B, t_end, s->stream), "batch_advance");
  if (scheme == SWE_SCHEME_EULER) s->cur = b;
  return SWE_OK;
}

int swe_batch_reset(swe_solver* s, double t0) {
  if (!s || s->n_inst <= 0) return fail(SWE_ERR_ARG, "not an ensemble solver");
  if (int rc = set_device(s)) return rc;
  const int n = s->n_inst;
  std::vector<double> t(size_t(n), t0), dt(size_t(n), 0.);
  std::vector<int> act(size_t(n), 1);
  std::vector<long long> steps(size_t(n), 0);
  SWE_CHECK(hipMemcpy(s->B.t, t.data(), sizeof(double) * size_t(n), hipMemcpyHostToDevice), "hipMemcpy");
  SWE_CHECK(hipMemcpy(s->B.dt, dt.data(), sizeof(double) * size_t(n), hipMemcpyHostToDevice), "hipMemcpy");
  SWE_CHECK(hipMemcpy(s->B.active, act.data(), sizeof(int) * size_t(n), hipMemcpyHostToDevice), "hipMemcpy");
  SWE_CHECK(hipMemcpy(s->B.steps, steps.data(), sizeof(long long) * size_t(n), hipMemcpyHostToDevice), "hipMemcpy");
  SWE_CHECK(hipMemcpy(s->B.n_active, &n, sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
  return SWE_OK;
}

// Cells and edges of the members still running: finished members at the
// two ends of the member list are left out of the launches.  Inside the
// window a finished member's state is carried along by the update (copied
// from the stage input); an Euler step swaps the two state buffers, so the
// cells that leave the window are copied into the other buffer as well,
// and both buffers then hold their final state whatever the later parity.
static int batch_window(swe_solver* s, const std::vector<int>& act, bool euler) {
  const int64_t oc0 = s->M.c0, oc1 = s->M.c1;
  int k0 = 0, k1 = s->n_inst - 1;
  while (k0 <= k1 && !act[size_t(k0)]) ++k0;
  while (k1 >= k0 && !act[size_t(k1)]) --k1;
  if (k0 > k1) {
    s->M.c0 = s->M.c1 = 0;
    s->M.e0 = s->M.e1 = 0;
  } else {
    s->M.c0 = s->cell_off[size_t(k0)];
    s->M.c1 = s->cell_off[size_t(k1) + 1];
    s->M.e0 = s->edge_off[size_t(k0)];
    s->M.e1 = s->edge_off[size_t(k1) + 1];
  }
  if (!euler) return SWE_OK;
  const int a = s->cur, b = a == 0 ? 1 : 0;
  auto copy = [&](int64_t c0, int64_t c1) -> int {
    if (c1 <= c0) return SWE_OK;
    if (s->lay_b != 0) {
      // blocked layouts: the whole blocks the cells lie in, one contiguous
      // range (cells of running members that share a block are copied too:
      // the other buffer's values of those are overwritten by their next
      // stage before anything reads them)
      const int64_t B = s->lay_b, k0 = c0 / B, k1 = (c1 + B - 1) / B;
      SWE_CHECK(hipMemcpyAsync(s->sbuf[b] + 3 * B * k0, s->sbuf[a] + 3 * B * k0, sizeof(double) * size_t(3 * B * (k1 - k0)),
                               hipMemcpyDeviceToDevice, s->stream), "hipMemcpy");
      return SWE_OK;
    }
    const size_t off = size_t(c0), bytes = sizeof(double) * size_t(c1 - c0);
    SWE_CHECK(hipMemcpyAsync(s->sw[b] + off, s->sw[a] + off, bytes, hipMemcpyDeviceToDevice, s->stream), "hipMemcpy");
    SWE_CHECK(hipMemcpyAsync(s->su[b] + off, s->su[a] + off, bytes, hipMemcpyDeviceToDevice, s->stream), "hipMemcpy");
    SWE_CHECK(hipMemcpyAsync(s->sv[b] + off, s->sv[a] + off, bytes, hipMemcpyDeviceToDevice, s->stream), "hipMemcpy");
    return SWE_OK;
  };
  if (s->M.c1 <= s->M.c0) return copy(oc0, oc1);
  if (int rc = copy(oc0, std::min(oc1, s->M.c0))) return rc;
  return copy(std::max(oc0, s->M.c1), oc1);
}

static int batch_capture(swe_solver* s, int scheme, double cfl, double t_end, int per_graph) {
  if (s->graph) { (void)hipGraphExecDestroy(s->graph); s->graph = nullptr; }
  const int cur0 = s->cur;
  hipGraph_t g;
  SWE_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  int rc = SWE_OK;
  for (int q = 0; q < per_graph && !rc; ++q) rc = batch_step(s, scheme, cfl, t_end);
  hipError_t ce = hipStreamEndCapture(s->stream, &g);
  s->cur = cur0;
  if (rc) return rc;
  if (ce != hipSuccess) return dev_fail(ce, "hipStreamEndCapture");
  hipError_t ie = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) return dev_fail(ie, "hipGraphInstantiate");
  s->graph_scheme = scheme;
  s->graph_cfl = cfl;
  s->graph_tend = t_end;
  s->graph_c0 = s->M.c0;
  s->graph_c1 = s->M.c1;
  s->graph_cur = cur0;
  return SWE_OK;
}

int swe_batch_run(swe_solver* s, int scheme, double cfl, double t_end, int64_t max_steps,
                  int use_graph, int64_t* steps_done, int64_t* n_active) {
  if (!s || s->n_inst <= 0) return fail(SWE_ERR_ARG, "not an ensemble solver");
  if (scheme < 0 || scheme > 2) return fail(SWE_ERR_ARG, "unknown scheme");
  if (!(cfl > 0.)) return fail(SWE_ERR_ARG, "ensembles run in CFL mode: cfl > 0");
  if (int rc = set_device(s)) return rc;
  // Euler ping-pongs between two state buffers: a graph holds two steps
  const int per_graph = (scheme == SWE_SCHEME_EULER) ? 2 : 1;
  std::vector<int> act_flags(size_t(s->n_inst));
  SWE_CHECK(hipMemcpyAsync(act_flags.data(), s->B.active, sizeof(int) * act_flags.size(),
                           hipMemcpyDeviceToHost, s->stream), "hipMemcpy");
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  const bool euler = scheme == SWE_SCHEME_EULER;
  if (int rc = batch_window(s, act_flags, euler)) return rc;
  int64_t done = 0;
  int act = 0;
  for (int f : act_flags) act += f;
  const int64_t chunk = 64;
  while (done < max_steps && act > 0) {
    if (use_graph) {
      // re-capture when the scheme changed, the window shrank (members
      // finished: at most one capture per check, every 64 steps) or the
      // state moved to the other buffer since the capture (an odd number of
      // Euler steps taken without the graph)
      const int64_t w = s->M.c1 - s->M.c0, wg = s->graph_c1 - s->graph_c0;
      if (s->graph == nullptr || s->graph_scheme != scheme || s->graph_cfl != cfl ||
          s->graph_tend != t_end || s->M.c0 < s->graph_c0 || s->M.c1 > s->graph_c1 || w < wg ||
          s->cur != s->graph_cur)
        if (int rc = batch_capture(s, scheme, cfl, t_end, per_graph)) return rc;
    }
    const int64_t k = std::min<int64_t>(chunk, max_steps - done);
    for (int64_t q = 0; q < k;) {
      if (use_graph && k - q >= per_graph) {
        SWE_CHECK(hipGraphLaunch(s->graph, s->stream), "hipGraphLaunch");
        q += per_graph;
      } else {  // without the graph, or the odd last step of an Euler graph
        if (int rc = batch_step(s, scheme, cfl, t_end)) return rc;
        q += 1;
      }
    }
    done += k;
    SWE_CHECK(hipMemcpyAsync(act_flags.data(), s->B.active, sizeof(int) * act_flags.size(),
                             hipMemcpyDeviceToHost, s->stream), "hipMemcpy");
    SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    act = 0;
    for (int f : act_flags) act += f;
    // a captured graph keeps its own window until the next capture; members
    // inside it that finished are skipped by the kernels
    if (int rc = batch_window(s, act_flags, euler)) return rc;
  }
  // restore the whole-mesh window for the single-mesh entry points
  s->M.c0 = 0;
  s->M.c1 = s->T;
  s->M.e0 = 0;
  s->M.e1 = s->E;
  if (steps_done) *steps_done = done;
  if (n_active) *n_active = act;
  return SWE_OK;
}

int swe_batch_info(swe_solver* s, int64_t* n, int64_t* cell_off, int64_t* edge_off,
                   int64_t* node_off) {
  if (!s || s->n_inst <= 0) return fail(SWE_ERR_ARG, "not an ensemble solver");
  if (n) *n = s->n_inst;
  const size_t m = size_t(s->n_inst) + 1;
  if (cell_off) std::memcpy(cell_off, s->cell_off.data(), sizeof(int64_t) * m);
  if (edge_off) std::memcpy(edge_off, s->edge_off.data(), sizeof(int64_t) * m);
  if (node_off) std::memcpy(node_off, s->node_off.data(), sizeof(int64_t) * m);
  return SWE_OK;
}

int swe_batch_time(swe_solver* s, double* t, int64_t* steps) {
  if (!s || s->n_inst <= 0) return fail(SWE_ERR_ARG, "not an ensemble solver");
  if (int rc = set_device(s)) return rc;
  const size_t n = size_t(s->n_inst);
  if (t) if (int rc = download(s, t, s->B.t, sizeof(double) * n)) return rc;
  if (steps) {
    std::vector<long long> st(n);
    if (int rc = download(s, st.data(), s->B.steps, sizeof(long long) * n)) return rc;
    for (size_t k = 0; k < n; ++k) steps[k] = st[k];
  }
  return SWE_OK;
}

int swe_time_run(swe_solver* s, int scheme, double dt, double cfl, int64_t nsteps, double* ms,
                 double* kernel_ms) {
  if (int rc = single_mesh(s)) return rc;
  if (!ms) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  hipEvent_t e0, e1;
  SWE_CHECK(hipEventCreate(&e0), "hipEventCreate");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(SWE_ERR_DEVICE, "hipEventCreate");
  }
  int rc = SWE_OK;
  if (!kernel_ms) {
    hipError_t e = hipEventRecord(e0, s->stream);
    if (e == hipSuccess) {
      rc = swe_run(s, scheme, dt, cfl, nsteps);
      e = hipEventRecord(e1, s->stream);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    if (e != hipSuccess && !rc) rc = dev_fail(e, "timing events");
    *ms = f;
  } else {
    // bracket each kernel of each stage with its own event pair: recon1,
    // recon2, then flux, drain, update or the fused kernel
    const int nk = 7;
    double acc[nk] = {0, 0, 0, 0, 0, 0, 0};
    int64_t cnt[nk] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<hipEvent_t> ev(2 * nk, nullptr);
    for (auto& x : ev)
      if (!rc && hipEventCreate(&x) != hipSuccess) rc = fail(SWE_ERR_DEVICE, "hipEventCreate");
    float total = 0.f;
    int nst = scheme == SWE_SCHEME_EULER ? 1 : (scheme == SWE_SCHEME_SSPRK2 ? 2 : 3);
    if (nsteps > s->dts_cap) {
      if (s->dts) { (void)hipStreamSynchronize(s->stream); (void)hipFree(s->dts); s->dts = nullptr; }
      SWE_CHECK(hipMalloc((void**)&s->dts, sizeof(double) * size_t(nsteps)), "hipMalloc");
      s->dts_cap = nsteps;
    }
    for (int64_t k = 0; k < nsteps && !rc; ++k) {
      int a = s->cur, b = (a + 1) % 3, c = (a + 2) % 3;
      for (int st = 0; st < nst && !rc; ++st) {
        int in, u0, out;
        double a0, a1, coef;
        if (st == 0) { in = a; u0 = a; out = b; a0 = 0; a1 = 0; coef = 1; }
        else if (scheme == SWE_SCHEME_SSPRK2) { in = b; u0 = a; out = a; a0 = .5; a1 = .5; coef = .5; }
        else if (st == 1) { in = b; u0 = a; out = c; a0 = .75; a1 = .25; coef = .25; }
        else { in = c; u0 = a; out = a; a0 = 1. / 3.; a1 = 2. / 3.; coef = 2. / 3.; }
        DevState S = state_of(s, in);
        StageArgs A{a0, a1, coef, dt, cfl, s->cor, s->rbits, st == 0 ? s->dts + k : nullptr, friction(s)};
        bool used[nk] = {false, false, false, false, false, false, false};
        hipError_t e = hipSuccess;
        auto timed = [&](int q, auto&& launch) {
          if (e == hipSuccess) e = hipEventRecord(ev[size_t(2 * q)], s->stream);
          if (e == hipSuccess) e = launch();
          if (e == hipSuccess) e = hipEventRecord(ev[size_t(2 * q + 1)], s->stream);
          used[q] = true;
        };
        timed(0, [&] { return launch_recon1(s->M, S, s->W, s->opts, s->stream); });
        timed(1, [&] { return launch_recon2(s->M, S, s->W, s->opts, s->stream); });
        if (use_fused(s, dt, st)) {
          timed(5, [&] {
            return launch_fused(s->M, s->tiles, S, state_of(s, u0), out_of(s, out), s->W, A, s->flux, s->ws,
                                s->opts, s->stream);
          });
          s->flux_stale = true;
          s->stale_in = in;
          s->stale_slot = st;
        } else {
          timed(2, [&] {
            return launch_flux(s->M, S, s->W, s->flux, s->ws, s->cor, s->rbits + st, s->flux_items, s->flux_block,
                               s->stream);
          });
          s->flux_stale = false;
          if (use_fused_after_flux(s)) {
            timed(6, [&] {
              return launch_fused(s->M, s->tiles, S, state_of(s, u0), out_of(s, out), s->W, A, s->flux, s->ws,
                                  s->opts, s->stream, false, true);
            });
          } else {
            timed(3, [&] { return launch_drain(s->M, S, s->W, s->opts, s->stream); });
            timed(4, [&] {
              return launch_update(s->M, S, state_of(s, u0), out_of(s, out), s->W, A, s->upd_items, s->upd_block,
                                   s->stream);
            });
          }
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) { rc = dev_fail(e, "kernel"); break; }
        for (int q = 0; q < nk; ++q) {
          if (!used[q]) continue;
          float f = 0.f;
          if (hipEventElapsedTime(&f, ev[size_t(2 * q)], ev[size_t(2 * q + 1)]) != hipSuccess) f = 0.f;
          acc[q] += f;
          cnt[q] += 1;
          total += f;
        }
        s->last_r_slot = st;
      }
      if (scheme == SWE_SCHEME_EULER) s->cur = b;
    }
    s->last_nsteps = nsteps;
    for (int q = 0; q < nk; ++q) kernel_ms[q] = cnt[q] ? acc[q] / double(cnt[q]) : 0.;
    *ms = total;
    for (auto& x : ev)
      if (x) (void)hipEventDestroy(x);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = dev_fail(e, "run");
  }
  return rc;
}

int swe_set_fused(swe_solver* s, int on, int* active) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  s->fused_on = on != 0;
  if (active) *active = (s->fused_ok && s->fused_on) ? 1 : 0;
  return SWE_OK;
}

int swe_set_tuning(swe_solver* s, int flux_items, int update_items) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (flux_items != 1 && flux_items != 2 && flux_items != 4)
    return fail(SWE_ERR_ARG, "flux_items must be 1, 2 or 4");
  if (update_items != 1 && update_items != 2) return fail(SWE_ERR_ARG, "update_items must be 1 or 2");
  if (s->flux_items != flux_items || s->upd_items != update_items) {
    s->flux_items = flux_items;
    s->upd_items = update_items;
    if (s->graph) {  // a captured ensemble step holds the old launches
      (void)hipGraphExecDestroy(s->graph);
      s->graph = nullptr;
      s->graph_scheme = -1;
    }
  }
  return SWE_OK;
}

int swe_get_tuning(const swe_solver* s, int* flux_items, int* update_items) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (flux_items) *flux_items = s->flux_items;
  if (update_items) *update_items = s->upd_items;
  return SWE_OK;
}

// Times every work-items-per-thread instantiation of the flux and update
// kernels on the current state (each launch is idempotent: the flux pass
// rewrites F and a CFL slot no step reads, the update writes a free state
// buffer), keeps the fastest, and reports the per-launch times in ms as
// {flux x1, x2, x4, update x1, x2}.  The state is left unchanged.
int swe_set_tuning_blocks(swe_solver* s, int flux_block, int update_block) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if ((flux_block != 256 && flux_block != 512) || (update_block != 256 && update_block != 512))
    return fail(SWE_ERR_ARG, "blocks must be 256 or 512 threads");
  if (s->flux_block != flux_block || s->upd_block != update_block) {
    s->flux_block = flux_block;
    s->upd_block = update_block;
    if (s->graph) {  // a captured ensemble step holds the old launches
      (void)hipGraphExecDestroy(s->graph);
      s->graph = nullptr;
      s->graph_scheme = -1;
    }
  }
  return SWE_OK;
}
int swe_get_tuning_blocks(const swe_solver* s, int* flux_block, int* update_block) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (flux_block) *flux_block = s->flux_block;
  if (update_block) *update_block = s->upd_block;
  return SWE_OK;
}
int swe_autotune(swe_solver* s, int reps, int* flux_items, int* update_items, double* times) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (reps < 1) return fail(SWE_ERR_ARG, "reps must be >= 1");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;  // the free buffer may hold a fused stage's input
  DevState S = state_of(s, s->cur);
  const int scratch = (s->cur + 1) % 3;
  unsigned long long* rslot = (s->n_inst > 0 ? s->B.rbits : s->rbits) + 3;
  SWE_CHECK(launch_recon1(s->M, S, s->W, s->opts, s->stream), "recon1");
  SWE_CHECK(launch_recon2(s->M, S, s->W, s->opts, s->stream), "recon2");
  hipEvent_t e0, e1;
  SWE_CHECK(hipEventCreate(&e0), "hipEventCreate");
  SWE_CHECK(hipEventCreate(&e1), "hipEventCreate");
  const int fi[3] = {1, 2, 4}, ui[2] = {1, 2}, bs[2] = {256, 512};
  double t[5][2];
  int rc = SWE_OK;
  // pass 0 only warms every variant up (clocks, caches), so that the
  // first one timed is not the one that pays for it; each variant keeps the
  // faster of passes 1 and 2
  for (int q = 0; q < 5; ++q) t[q][0] = t[q][1] = 1e300;
  auto timed = [&](auto&& launch, double& out, bool keep) {
    hipError_t e = hipEventRecord(e0, s->stream);
    for (int q = 0; q < reps && e == hipSuccess; ++q) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    if (e != hipSuccess) return dev_fail(e, "autotune");
    if (keep && double(f) / reps < out) out = double(f) / reps;
    return SWE_OK;
  };
  // Euler-form stage with a small fixed step into the free buffer
  StageArgs A{0., 0., 1., 1e-6, 0., s->cor, rslot, nullptr, friction(s)};
  for (int pass = 0; pass < 3 && !rc; ++pass) {
    for (int q = 0; q < 3 && !rc; ++q)
      for (int b = 0; b < 2 && !rc; ++b)
        rc = timed([&] { return launch_flux(s->M, S, s->W, s->flux, s->ws, s->cor, rslot, fi[q], bs[b], s->stream); },
                   t[q][b], pass > 0);
    if (!rc) {
      hipError_t e = launch_drain(s->M, S, s->W, s->opts, s->stream);
      if (e != hipSuccess) rc = dev_fail(e, "drain");
    }
    for (int q = 0; q < 2 && !rc; ++q)
      for (int b = 0; b < 2 && !rc; ++b)
        rc = timed([&] {
          return launch_update(s->M, S, S, out_of(s, scratch), s->W, A, ui[q], bs[b], s->stream);
        }, t[3 + q][b], pass > 0);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  int bf = 0, bfb = 0, bu = 0, bub = 0;
  for (int q = 0; q < 3; ++q)
    for (int b = 0; b < 2; ++b)
      if (t[q][b] < t[bf][bfb]) { bf = q; bfb = b; }
  for (int q = 0; q < 2; ++q)
    for (int b = 0; b < 2; ++b)
      if (t[3 + q][b] < t[3 + bu][bub]) { bu = q; bub = b; }
  if (int r2 = swe_set_tuning(s, fi[bf], ui[bu])) return r2;
  if (int r2 = swe_set_tuning_blocks(s, bs[bfb], bs[bub])) return r2;
  if (s->n_inst == 0) s->last_r_slot = 3;
  if (flux_items) *flux_items = fi[bf];
  if (update_items) *update_items = ui[bu];
  if (times) for (int q = 0; q < 5; ++q) times[q] = t[q][0] < t[q][1] ? t[q][0] : t[q][1];
  return SWE_OK;
}

// The state layout: the current state is rewritten into a free buffer in
// the new layout by a device kernel and becomes current; the other buffers
// take the new layout's pointers (their contents are stage scratch).
int swe_set_layout(swe_solver* s, int block) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (block < 0 || block > kLayoutMaxBlock || (block & (block - 1)) != 0)
    return fail(SWE_ERR_ARG, "layout block: 0 (SoA) or a power of two up to 64 cells");
  if (block == s->lay_b) return SWE_OK;
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;  // it reads the stale stage input in the old layout
  const DevState src = state_of(s, s->cur);
  const int to = (s->cur + 1) % 3;
  s->lay_b = block;
  layout_ptrs(s);
  SWE_CHECK(launch_relayout(src, out_of(s, to), s->T, s->stream), "relayout");
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  s->cur = to;
  if (s->graph) {  // captured launches hold the old pointers
    (void)hipGraphExecDestroy(s->graph);
    s->graph = nullptr;
    s->graph_scheme = -1;
  }
  return SWE_OK;
}
int swe_get_layout(const swe_solver* s, int* block) {
  if (!s || !block) return fail(SWE_ERR_ARG, "null argument");
  *block = s->lay_b;
  return SWE_OK;
}

// The layout dimension of the autotuner: for each candidate layout the
// work-item and workgroup variants are tuned (swe_autotune), then one stage
// of each kind the solver runs is timed on the live state with them: the
// first stage of a CFL step (reconstruction, flux pass, then the tile kernel
// reading the fluxes, or the separate draining and update kernels where no
// tiles are built), and for a solver that fuses its later stages also the
// reconstruction + fused stage.  The layout (with its variants) of the least
// cost, first stage + (fused or first-stage kind) stage (an SSP-RK2 step in
// CFL mode), is kept.
int swe_autotune_layout(swe_solver* s, int reps, int* block, double* cost_ms) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (reps < 1) return fail(SWE_ERR_ARG, "reps must be >= 1");
  if (int rc = set_device(s)) return rc;
  constexpr int kLayoutCandidates = SWE_LAYOUT_CANDIDATES;
  const int cand[kLayoutCandidates] = {0, 64, 8, 1};
  double cost[kLayoutCandidates];
  int items[kLayoutCandidates][2], blocks[kLayoutCandidates][2];
  hipEvent_t e0, e1;
  SWE_CHECK(hipEventCreate(&e0), "hipEventCreate");
  SWE_CHECK(hipEventCreate(&e1), "hipEventCreate");
  int rc = SWE_OK;
  const bool fused = use_fused(s, 1e-6, 1);
  unsigned long long* rslot = (s->n_inst > 0 ? s->B.rbits : s->rbits) + 3;
  for (int c = 0; c < kLayoutCandidates && !rc; ++c) {
    rc = swe_set_layout(s, cand[c]);
    if (!rc) rc = swe_autotune(s, reps, &items[c][0], &items[c][1], nullptr);
    if (!rc) rc = swe_get_tuning_blocks(s, &blocks[c][0], &blocks[c][1]);
    if (rc) break;
    const DevState S = state_of(s, s->cur);
    const DevOut O = out_of(s, (s->cur + 1) % 3);
    const StageArgs A{0., 0., 1., 1e-6, 0., s->cor, rslot, nullptr, friction(s)};
    auto sep = [&]() -> hipError_t {
      hipError_t e = launch_recon1(s->M, S, s->W, s->opts, s->stream);
      if (e == hipSuccess) e = launch_recon2(s->M, S, s->W, s->opts, s->stream);
      if (e == hipSuccess)
        e = launch_flux(s->M, S, s->W, s->flux, s->ws, s->cor, rslot, s->flux_items, s->flux_block, s->stream);
      if (e == hipSuccess && fused)
        e = launch_fused(s->M, s->tiles, S, S, O, s->W, A, s->flux, s->ws, s->opts, s->stream, false, true);
      if (e == hipSuccess && !fused) e = launch_drain(s->M, S, s->W, s->opts, s->stream);
      if (e == hipSuccess && !fused) e = launch_update(s->M, S, S, O, s->W, A, s->upd_items, s->upd_block, s->stream);
      return e;
    };
    auto fus = [&]() -> hipError_t {
      hipError_t e = launch_recon1(s->M, S, s->W, s->opts, s->stream);
      if (e == hipSuccess) e = launch_recon2(s->M, S, s->W, s->opts, s->stream);
      if (e == hipSuccess) e = launch_fused(s->M, s->tiles, S, S, O, s->W, A, s->flux, s->ws, s->opts, s->stream);
      return e;
    };
    double t[2] = {1e300, 1e300};
    for (int pass = 0; pass < 3 && !rc; ++pass)  // pass 0 warms up
      for (int k = 0; k < (fused ? 2 : 1) && !rc; ++k) {
        hipError_t e = hipEventRecord(e0, s->stream);
        for (int q = 0; q < reps && e == hipSuccess; ++q) e = k == 0 ? sep() : fus();
        if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float f = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
        if (e != hipSuccess) rc = dev_fail(e, "autotune layout");
        else if (pass > 0 && double(f) / reps < t[k]) t[k] = double(f) / reps;
      }
    cost[c] = t[0] + (fused ? t[1] : t[0]);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  int best = 0;
  for (int c = 1; c < kLayoutCandidates; ++c)
    if (cost[c] < cost[best]) best = c;
  if (int r2 = swe_set_layout(s, cand[best])) return r2;
  if (int r2 = swe_set_tuning(s, items[best][0], items[best][1])) return r2;
  if (int r2 = swe_set_tuning_blocks(s, blocks[best][0], blocks[best][1])) return r2;
  if (block) *block = cand[best];
  if (cost_ms) for (int c = 0; c < kLayoutCandidates; ++c) cost_ms[c] = cost[c];
  return SWE_OK;
}

// ComputeIntegrals of the current state: reconstruction, per-cell terms
// and a fixed-order reduction (identical run to run).
int swe_integrals(swe_solver* s, double* out3, double* per_cell) {
  if (int rc = single_mesh(s)) return rc;
  if (!out3) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  DevState S = state_of(s, s->cur);
  SWE_CHECK(launch_recon1(s->M, S, s->W, s->opts, s->stream), "recon1");
  SWE_CHECK(launch_recon2(s->M, S, s->W, s->opts, s->stream), "recon2");
  const int64_t nb = integrals_blocks(s->M);
  double *d = nullptr, *pc = nullptr;
  SWE_CHECK(hipMalloc((void**)&d, sizeof(double) * size_t(3 * nb + 3)), "hipMalloc");
  hipError_t e = hipSuccess;
  if (per_cell) e = hipMalloc((void**)&pc, sizeof(double) * size_t(3 * std::max<int64_t>(s->T, 1)));
  if (e == hipSuccess) e = launch_integrals(s->M, S, s->W, d + 3, pc, d, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out3, d, 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess && per_cell)
    e = hipMemcpyAsync(per_cell, pc, sizeof(double) * size_t(3 * s->T), hipMemcpyDeviceToHost, s->stream);
  hipError_t e2 = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (pc) (void)hipFree(pc);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return dev_fail(e, "integrals");
  return SWE_OK;
}

// swe_integrals without allocation or host synchronisation: the block
// partials live in the device's static scratch array, the results go to
// device memory on the solver's stream.
int swe_integrals_noalloc(swe_solver* s, double* out3_dev, double* per_cell_dev) {
  if (int rc = single_mesh(s)) return rc;
  if (!out3_dev) return fail(SWE_ERR_ARG, "null argument");
  if (!s->scratch) return fail(SWE_ERR_DEVICE, "no scratch array on this device");
  const int64_t nb = integrals_blocks(s->M);
  if (nb > kScratchBlocks)
    return fail(SWE_ERR_ARG, "mesh exceeds the allocation-free limit of " +
                                 std::to_string(kScratchBlocks * 256) + " cells");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  DevState S = state_of(s, s->cur);
  SWE_CHECK(launch_recon1(s->M, S, s->W, s->opts, s->stream), "recon1");
  SWE_CHECK(launch_recon2(s->M, S, s->W, s->opts, s->stream), "recon2");
  SWE_CHECK(launch_integrals(s->M, S, s->W, s->scratch, per_cell_dev, out3_dev, s->stream), "integrals");
  return SWE_OK;
}

// Final reconstruction of every cell after swe_compute_interface_values:
// origin level w0 at the centroid and gradient (gx, gy) (any may be null).
int swe_get_muscl(swe_solver* s, double* w0, double* gx, double* gy) {
  if (!s) return fail(SWE_ERR_ARG, "null solver");
  if (int rc = set_device(s)) return rc;
  const size_t b = sizeof(double) * size_t(s->T);
  if (w0) if (int rc = download(s, w0, s->W.mw, b)) return rc;
  if (gx) if (int rc = download(s, gx, s->W.mgx, b)) return rc;
  if (gy) if (int rc = download(s, gy, s->W.mgy, b)) return rc;
  return SWE_OK;
}

// Health of the current state: out5 = (cells with a non-finite value,
// min h, max h, max |u|, max |u| + sqrt(h)) over all cells (wet cells for
// the speeds).  Exact and identical run to run.
int swe_state_stats(swe_solver* s, double* out5) {
  if (!s || !out5) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  const int64_t nb = integrals_blocks(s->M);
  double* d = nullptr;
  SWE_CHECK(hipMalloc((void**)&d, sizeof(double) * size_t(5 * nb + 5)), "hipMalloc");
  hipError_t e = launch_stats(s->M, state_of(s, s->cur), d + 5, d, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out5, d, 5 * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  hipError_t e2 = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return dev_fail(e, "state_stats");
  return SWE_OK;
}

// ------------------------------------------------------------- adjoint --

namespace {

// optional timing of the adjoint and tangent sweeps: events that cannot be
// created or recorded leave the reported times at zero and do not affect
// the computation
struct SweepTimer {
  hipEvent_t e[3] = {nullptr, nullptr, nullptr};
  SweepTimer() {
    for (auto& x : e)
      if (hipEventCreate(&x) != hipSuccess) x = nullptr;
  }
  ~SweepTimer() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
  void mark(int k, hipStream_t st) {
    if (e[k] && hipEventRecord(e[k], st) != hipSuccess) {
      (void)hipEventDestroy(e[k]);
      e[k] = nullptr;
    }
  }
  float ms(int a, int b) const {  // after a synchronisation of the stream
    float f = 0.f;
    if (!e[a] || !e[b] || hipEventElapsedTime(&f, e[a], e[b]) != hipSuccess) f = 0.f;
    return f;
  }
};

// a state held in one 3T allocation (w, u, v)
struct Buf3 {
  double* p = nullptr;
};
DevState st3(const Buf3& b, int64_t T) { return DevState{b.p, b.p + T, b.p + 2 * T}; }
DevOut out3(const Buf3& b, int64_t T) { return DevOut{b.p, b.p + T, b.p + 2 * T}; }

hipError_t copy_state(const DevState& from, const DevOut& to, int64_t T, hipStream_t st) {
  const size_t n = sizeof(double) * size_t(T);
  hipError_t e = hipMemcpyAsync(to.w, from.w, n, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(to.u, from.u, n, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(to.v, from.v, n, hipMemcpyDeviceToDevice, st);
  return e;
}

// Shu-Osher stages of a scheme: (a0, a1, coef); stage q reads the previous
// stage's output (the step input for q = 0) and the step input as U0
struct StageCoef {
  double a0, a1, coef;
};
int stage_list(int scheme, StageCoef* out) {
  if (scheme == SWE_SCHEME_EULER) { out[0] = {0., 0., 1.}; return 1; }
  if (scheme == SWE_SCHEME_SSPRK2) { out[0] = {0., 0., 1.}; out[1] = {0.5, 0.5, 0.5}; return 2; }
  out[0] = {0., 0., 1.};
  out[1] = {0.75, 0.25, 0.25};
  out[2] = {1. / 3., 2. / 3., 2. / 3.};
  return 3;
}

}  // namespace

// Adjoint of nsteps steps (no reference counterpart: the differentiation of
// the full step).  Starting from the current state U_0, runs the steps with
// step sizes dt (> 0, every step) or dts[n] (e.g. a CFL run's sizes, then
// held fixed), storing U_n every checkpoint_every steps; leaves U_N as the
// current state.  Each segment between checkpoints is then recomputed from
// its checkpoint with every stage's intermediates (MUSCL, fluxes, draining
// dt) kept in device memory, and final_bar (T x 3 row-major adjoint of U_N's
// w, u, v) is propagated back through its stages into state0_bar (T x 3,
// adjoint of U_0) and z_bar (V, adjoint of the node bathymetry; may be
// null).  The last segment's recomputation is the forward run itself, so
// checkpoint_every >= nsteps runs the steps once.  ms (optional, 2
// doubles): forward sweep and reverse sweep (recomputation included)
// milliseconds.
int swe_adjoint_run(swe_solver* s, int scheme, double dt, const double* dts, int64_t nsteps,
                    int64_t checkpoint_every, const double* final_bar, double* state0_bar, double* z_bar,
                    double* ms) {
  if (int rc = single_mesh(s)) return rc;
  if (!final_bar || !state0_bar) return fail(SWE_ERR_ARG, "null argument");
  if (scheme < 0 || scheme > 2) return fail(SWE_ERR_ARG, "unknown scheme");
  if (nsteps < 1) return fail(SWE_ERR_ARG, "nsteps must be >= 1");
  if (!dts && !(dt > 0.)) return fail(SWE_ERR_ARG, "the adjoint needs fixed step sizes (dt > 0 or dts)");
  if (dts)
    for (int64_t n = 0; n < nsteps; ++n)
      if (!(dts[n] > 0.)) return fail(SWE_ERR_ARG, "step sizes must be positive");
  if (s->M.pe) return fail(SWE_ERR_ARG, "periodic boundaries are not differentiated");
  if (s->lay_b != 0) return fail(SWE_ERR_ARG, "the adjoint and tangent runs need the SoA state layout (swe_set_layout(s, 0))");
  if (s->opts & (SWE_OPT_HEAD_GRADIENT | SWE_OPT_HEAD_PARTWET2 | SWE_OPT_EXACT_BALANCE))
    return fail(SWE_ERR_ARG, "the adjoint does not support the HEAD-compatibility or exact-balance options");
  const int64_t K = checkpoint_every > 0 ? std::min(checkpoint_every, nsteps) : nsteps;
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  const int64_t T = s->T, E = s->E, V = s->V;
  StageCoef sc[3];
  const int ns = stage_list(scheme, sc);
  const int64_t nck = (nsteps + K - 1) / K;
  std::vector<void*> mem;
  auto alloc = [&](auto** p, int64_t bytes) {
    hipError_t e = hipMalloc((void**)p, size_t(std::max<int64_t>(bytes, 8)));
    if (e == hipSuccess) mem.push_back((void*)*p);
    return e;
  };
  auto release = [&]() {  // cleanup: errors are reported where they occur
    (void)hipStreamSynchronize(s->stream);
    for (void* p : mem) (void)hipFree(p);
  };
  const int64_t S3 = 3 * T * int64_t(sizeof(double));
  // per-stage work slot: mw, mgx, mgy, wrec1, dtd (T doubles each), f0, f1,
  // f2 (E each), pw_list (T ints), cls (T bytes), pw_count
  const int64_t slot_bytes = ((5 * T + 3 * E) * 8 + 4 * T + T + 64 + 255) / 256 * 256;
  double *ck = nullptr, *seg = nullptr, *stg = nullptr, *bar_a = nullptr, *bar_b = nullptr;
  double *zb = nullptr, *work = nullptr;
  char* slots = nullptr;
  unsigned int* cnt = nullptr;
  std::vector<unsigned int> hcnt(static_cast<size_t>(K * ns), kPwUnknown);
  hipError_t e = alloc(&ck, nck * S3);
  if (e == hipSuccess) e = alloc(&seg, K * S3);
  if (e == hipSuccess) e = alloc(&stg, K * std::max(ns - 1, 1) * S3);
  if (e == hipSuccess) e = alloc(&slots, K * ns * slot_bytes);
  if (e == hipSuccess) e = alloc(&bar_a, S3);
  if (e == hipSuccess) e = alloc(&bar_b, S3);
  if (e == hipSuccess) e = alloc(&zb, V * 8);
  // DevAdj work: rb 3T, dtd_own T, dtd_nb 3T, m_cell 3T, sl_edge 6E, drain_h T,
  // w_nb 3T, zc 3T, u0bar 3T, pw_own 4T, pw1_jac 4T, pw_w23bar T, p1bar 3T, zn_acc V,
  // zopp_acc 2E; back 3T and eloc 2E bytes
  const int64_t nwork = 3 * T + T + 3 * T + 3 * T + 6 * E + T + 3 * T + 3 * T + 3 * T + 4 * T + 4 * T + T + 3 * T + V + 2 * E;
  if (e == hipSuccess) e = alloc(&work, nwork * 8 + ((3 * T + 2 * E + 255) / 256) * 256);
  if (e == hipSuccess) e = alloc(&cnt, K * ns * int64_t(sizeof(unsigned int)));
  if (e != hipSuccess) { release(); return dev_fail(e, "hipMalloc (adjoint)"); }
  auto slot_work = [&](int64_t idx) {
    char* b = slots + idx * slot_bytes;
    DevWork W{};
    double* d = (double*)b;
    W.mw = d; d += T;
    W.mgx = d; d += T;
    W.mgy = d; d += T;
    W.wrec1 = d; d += T;
    W.dtd = d; d += T;
    W.f0 = d; d += E;
    W.f1 = d; d += E;
    W.f2 = d; d += E;
    int* ip = (int*)d;
    W.pw_list = ip; ip += T;
    W.cls = (int8_t*)ip;
    W.pw_count = (unsigned int*)((((uintptr_t)(W.cls + T)) + 63) / 64 * 64);
    return W;
  };
  DevAdj J{};
  {
    double* q = work;
    J.rb = q; q += 3 * T;
    J.dtd_own = q; q += T;
    J.dtd_nb = q; q += 3 * T;
    J.m_cell = q; q += 3 * T;
    J.sl_edge = q; q += 6 * E;
    J.drain_h = q; q += T;
    J.w_nb = q; q += 3 * T;
    J.zc = q; q += 3 * T;
    J.u0bar = q; q += 3 * T;
    J.pw_own = q; q += 4 * T;
    J.pw1_jac = q; q += 4 * T;
    J.pw_w23bar = q; q += T;
    J.p1bar = q; q += 3 * T;
    J.zn_acc = q; q += V;
    J.zopp_acc = q; q += 2 * E;
    J.back = (int8_t*)q;
    J.eloc = J.back + 3 * T;
  }
  auto dtn = [&](int64_t n) { return dts ? dts[n] : dt; };
  SweepTimer tm;
  // forward sweep: checkpoints up to the start of the last segment
  tm.mark(0, s->stream);
  int rc = SWE_OK;
  const int64_t n_last = (nck - 1) * K;
  for (int64_t n = 0; n < n_last && !rc; ++n) {
    if (n % K == 0)
      if ((e = copy_state(state_of(s, s->cur), out3(Buf3{ck + (n / K) * 3 * T}, T), T, s->stream)) != hipSuccess)
        rc = dev_fail(e, "copy");
    if (!rc) rc = run_step(s, scheme, dtn(n), 0., nullptr);
  }
  if (!rc && (e = copy_state(state_of(s, s->cur), out3(Buf3{ck + (nck - 1) * 3 * T}, T), T, s->stream)) != hipSuccess)
    rc = dev_fail(e, "copy");
  tm.mark(1, s->stream);
  if (!rc) {
    std::vector<double> h(static_cast<size_t>(3 * T));
    for (int64_t i = 0; i < T; ++i)
      for (int q = 0; q < 3; ++q) h[size_t(q * T + i)] = final_bar[3 * i + q];
    if ((e = hipMemcpyAsync(bar_a, h.data(), size_t(S3), hipMemcpyHostToDevice, s->stream)) != hipSuccess ||
        (e = hipMemsetAsync(J.zc, 0, sizeof(double) * size_t(3 * T), s->stream)) != hipSuccess ||
        (e = hipMemsetAsync(J.p1bar, 0, sizeof(double) * size_t(3 * T), s->stream)) != hipSuccess ||
        (e = hipMemsetAsync(J.zn_acc, 0, sizeof(double) * size_t(std::max<int64_t>(V, 1)), s->stream)) != hipSuccess ||
        (e = hipMemsetAsync(J.zopp_acc, 0, sizeof(double) * size_t(std::max<int64_t>(2 * E, 1)), s->stream)) != hipSuccess ||
        (e = hipMemsetAsync(zb, 0, sizeof(double) * size_t(std::max<int64_t>(V, 1)), s->stream)) != hipSuccess ||
        (e = launch_adjoint_tables(s->M, J, s->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(s->stream)) != hipSuccess)
      rc = dev_fail(e, "adjoint setup");
  }
  double* ob = bar_a;
  double* sb = bar_b;
  const double* wpend = nullptr;
  SweepTimer rt;  // each segment's recomputation, read after its synchronisation
  double recompute_ms = 0.;
  for (int64_t g = nck - 1; g >= 0 && !rc; --g) {
    const int64_t n0 = g * K, n1 = std::min(nsteps, n0 + K);
    // recompute the segment keeping step inputs, stage outputs and every
    // stage's intermediates; the last segment also produces U_N
    rt.mark(0, s->stream);
    e = copy_state(st3(Buf3{ck + g * 3 * T}, T), out3(Buf3{seg}, T), T, s->stream);
    // the slots' pass-2 counters, zeroed in one launch
    const int nslot = int((n1 - n0) * ns);
    if (e == hipSuccess) e = launch_slot_counts(slot_work(0).pw_count, slot_bytes, nslot, nullptr, s->stream);
    for (int64_t n = n0; n < n1 && e == hipSuccess; ++n) {
      const int64_t m = n - n0;
      const DevState Un = st3(Buf3{seg + m * 3 * T}, T);
      DevState Sq = Un;
      for (int q = 0; q < ns && e == hipSuccess; ++q) {
        DevOut O;
        if (q < ns - 1) O = out3(Buf3{stg + (m * (ns - 1) + q) * 3 * T}, T);
        else if (n + 1 < n1) O = out3(Buf3{seg + (m + 1) * 3 * T}, T);
        else O = g == nck - 1 ? out_of(s, (s->cur + 1) % 3) : out3(Buf3{sb}, T);  // U_N, or unused
        const DevWork Wq = slot_work(m * ns + q);
        e = launch_recon1(s->M, Sq, Wq, s->opts, s->stream, false);
        if (e == hipSuccess) e = launch_recon2(s->M, Sq, Wq, s->opts, s->stream);
        StageArgs A{sc[q].a0, sc[q].a1, sc[q].coef, dtn(n), 0., s->cor, s->rbits, nullptr, friction(s)};
        if (use_fused(s, dtn(n), q) && T < kFusedRecomputeMaxCells) {
          // one kernel: the stage output, and the fluxes and draining time
          // steps the reverse stage reads, stored into the slot (on meshes
          // below 2^20 cells, whose stages are launch-bound: above, the
          // fused kernel's edge-scattered flux stores cost more than the two
          // launches they save, 4 % of the reverse sweep at 8M cells)
          if (e == hipSuccess)
            e = launch_fused(s->M, s->tiles, Sq, Un, O, Wq, A, s->flux, s->ws, s->opts, s->stream, true);
        } else {
          if (e == hipSuccess)  // fixed dt: the CFL reduction is not read
            e = launch_flux(s->M, Sq, Wq, s->flux, s->ws, s->cor, s->rbits + 3, s->flux_items, s->flux_block, s->stream,
                            false);
          if (use_fused_after_flux(s)) {  // draining (stored) and update in one tile kernel
            if (e == hipSuccess)
              e = launch_fused(s->M, s->tiles, Sq, Un, O, Wq, A, s->flux, s->ws, s->opts, s->stream, true, true);
          } else {
            if (e == hipSuccess) e = launch_drain(s->M, Sq, Wq, s->opts, s->stream);
            if (e == hipSuccess) e = launch_update(s->M, Sq, Un, O, Wq, A, s->upd_items, s->upd_block, s->stream);
          }
        }
        Sq = DevState{O.w, O.u, O.v, O.mask};
      }
    }
    rt.mark(1, s->stream);
    // the lengths of the segment's pass-2 lists, for the partially wet
    // kernels of the reverse stages (one synchronisation per segment)
    if (e == hipSuccess) e = launch_slot_counts(slot_work(0).pw_count, slot_bytes, nslot, cnt, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hcnt.data(), cnt, sizeof(unsigned int) * size_t(nslot), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) { rc = dev_fail(e, "adjoint recomputation"); break; }
    recompute_ms += rt.ms(0, 1);
    if (g == nck - 1) s->cur = (s->cur + 1) % 3;  // U_N is the current state
    // reverse the steps of the segment: ob holds the adjoint of the stage
    // output, its w part owing the neighbour terms wpend of the stage after
    for (int64_t n = n1 - 1; n >= n0 && !rc; --n) {
      const int64_t m = n - n0;
      const DevState Un = st3(Buf3{seg + m * 3 * T}, T);
      for (int q = ns - 1; q >= 0 && e == hipSuccess; --q) {
        const DevState Sq = q == 0 ? Un : st3(Buf3{stg + (m * (ns - 1) + q - 1) * 3 * T}, T);
        J.obar = ob;
        J.sbar = sb;
        StageArgs A{sc[q].a0, sc[q].a1, sc[q].coef, dtn(n), 0., s->cor, s->rbits, nullptr, friction(s)};
        const int u0mode = (q < ns - 1 ? kU0Accumulate : 0) | (q == 0 && ns > 1 ? kU0Into : 0);
        e = launch_adjoint_stage(s->M, Sq, Un, slot_work(m * ns + q), A, J, s->flux, s->ws, s->opts, wpend,
                                 u0mode, hcnt[size_t(m * ns + q)], s->stream);
        wpend = J.w_nb;
        std::swap(ob, sb);
      }
      if (e != hipSuccess) rc = dev_fail(e, "adjoint stage");
    }
  }
  J.sbar = ob;  // the U_0 adjoint, its neighbour terms gathered here
  if (!rc && (e = launch_adjoint_finish(s->M, J, z_bar ? zb : nullptr, s->stream)) != hipSuccess)
    rc = dev_fail(e, "adjoint finish");
  tm.mark(2, s->stream);
  if (!rc) {
    std::vector<double> h(static_cast<size_t>(3 * T));
    if ((e = hipMemcpyAsync(h.data(), ob, size_t(S3), hipMemcpyDeviceToHost, s->stream)) != hipSuccess ||
        (z_bar && (e = hipMemcpyAsync(z_bar, zb, sizeof(double) * size_t(V), hipMemcpyDeviceToHost, s->stream)) != hipSuccess) ||
        (e = hipStreamSynchronize(s->stream)) != hipSuccess) {
      rc = dev_fail(e, "adjoint download");
    } else {
      for (int64_t i = 0; i < T; ++i)
        for (int q = 0; q < 3; ++q) state0_bar[3 * i + q] = h[size_t(q * T + i)];
      if (ms) {
        ms[0] = tm.ms(0, 1);
        ms[1] = tm.ms(1, 2);
      }
      s->adj_ms[0] = tm.ms(0, 1);
      s->adj_ms[1] = recompute_ms;
      s->adj_ms[2] = tm.ms(1, 2) - recompute_ms;
    }
  }
  release();
  return rc;
}

int swe_adjoint_timing(const swe_solver* s, double* out3) {
  if (!s || !out3) return fail(SWE_ERR_ARG, "null argument");
  for (int k = 0; k < 3; ++k) out3[k] = s->adj_ms[k];
  return SWE_OK;
}

// Tangent-linear run (no reference counterpart): the directional
// derivative of nsteps steps (fixed sizes dt or dts[n]) along the initial
// state tangent state0_dot (T x 3 row-major) and the node bathymetry tangent
// z_dot (V, may be null: zero), by the stage kernels built on dual numbers.
// The values advance as the forward step's, bit for bit, and U_N is the
// current state on return; final_dot receives the tangent of U_N (T x 3).
// ms (optional): milliseconds.
int swe_tangent_run(swe_solver* s, int scheme, double dt, const double* dts, int64_t nsteps,
                    const double* state0_dot, const double* z_dot, double* final_dot, double* ms) {
  if (int rc = single_mesh(s)) return rc;
  if (!state0_dot || !final_dot) return fail(SWE_ERR_ARG, "null argument");
  if (scheme < 0 || scheme > 2) return fail(SWE_ERR_ARG, "unknown scheme");
  if (nsteps < 1) return fail(SWE_ERR_ARG, "nsteps must be >= 1");
  if (!dts && !(dt > 0.)) return fail(SWE_ERR_ARG, "the tangent run needs fixed step sizes (dt > 0 or dts)");
  if (dts)
    for (int64_t n = 0; n < nsteps; ++n)
      if (!(dts[n] > 0.)) return fail(SWE_ERR_ARG, "step sizes must be positive");
  if (s->M.pe) return fail(SWE_ERR_ARG, "periodic boundaries are not differentiated");
  if (s->lay_b != 0) return fail(SWE_ERR_ARG, "the adjoint and tangent runs need the SoA state layout (swe_set_layout(s, 0))");
  if (s->opts & (SWE_OPT_HEAD_GRADIENT | SWE_OPT_HEAD_PARTWET2 | SWE_OPT_EXACT_BALANCE))
    return fail(SWE_ERR_ARG, "the tangent run does not support the HEAD-compatibility or exact-balance options");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  const int64_t T = s->T, E = s->E, V = s->V;
  // tangent states (three, rotating like the values), node z tangent,
  // intermediates
  const int64_t n8 = 9 * T + V + 5 * T + 3 * E;
  double* buf = nullptr;
  SWE_CHECK(hipMalloc((void**)&buf, sizeof(double) * size_t(n8)), "hipMalloc (tangent)");
  double* q = buf;
  double* td[3];
  for (int k = 0; k < 3; ++k) { td[k] = q; q += 3 * T; }
  double* zd = q; q += V;
  DevTan Q{};
  Q.zd = zd;
  Q.mw = q; q += T;
  Q.mgx = q; q += T;
  Q.mgy = q; q += T;
  Q.wrec1 = q; q += T;
  Q.dtd = q; q += T;
  Q.f0 = q; q += E;
  Q.f1 = q; q += E;
  Q.f2 = q; q += E;
  auto tst = [&](int k) { return DevState{td[k], td[k] + T, td[k] + 2 * T}; };
  auto tout = [&](int k) { return DevOut{td[k], td[k] + T, td[k] + 2 * T}; };
  std::vector<double> h(static_cast<size_t>(3 * T));
  for (int64_t i = 0; i < T; ++i)
    for (int c = 0; c < 3; ++c) h[size_t(c * T + i)] = state0_dot[3 * i + c];
  std::vector<double> hz(static_cast<size_t>(std::max<int64_t>(V, 1)), 0.);
  if (z_dot) std::copy(z_dot, z_dot + V, hz.begin());
  int rc = SWE_OK;
  hipError_t e = hipMemcpyAsync(td[s->cur], h.data(), sizeof(double) * size_t(3 * T), hipMemcpyHostToDevice,
                                s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(zd, hz.data(), sizeof(double) * size_t(V), hipMemcpyHostToDevice, s->stream);
  StageCoef sc[3];
  const int ns = stage_list(scheme, sc);
  SweepTimer tm;
  tm.mark(0, s->stream);
  for (int64_t n = 0; n < nsteps && e == hipSuccess; ++n) {
    // the slots of run_step: stage k reads the previous stage's output (the
    // step input a for k = 0) and writes outs[k]; U0 is slot a
    const int a = s->cur, b = (a + 1) % 3, c = (a + 2) % 3;
    const int outs[3] = {b, ns == 2 ? a : c, a};
    for (int k = 0; k < ns && e == hipSuccess; ++k) {
      const int in = k == 0 ? a : outs[k - 1];
      StageArgs A{sc[k].a0, sc[k].a1, sc[k].coef, dts ? dts[n] : dt, 0., s->cor, s->rbits, nullptr, friction(s)};
      e = launch_tangent_stage(s->M, state_of(s, in), tst(in), state_of(s, a), tst(a), out_of(s, outs[k]),
                               tout(outs[k]), s->W, Q, A, s->flux, s->ws, s->opts, s->stream);
    }
    s->cur = ns == 1 ? b : a;
  }
  tm.mark(1, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), td[s->cur], sizeof(double) * size_t(3 * T), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    rc = dev_fail(e, "tangent run");
  } else {
    for (int64_t i = 0; i < T; ++i)
      for (int c = 0; c < 3; ++c) final_dot[3 * i + c] = h[size_t(c * T + i)];
    if (ms) *ms = tm.ms(0, 1);
  }
  (void)hipStreamSynchronize(s->stream);  // cleanup: errors were reported above
  (void)hipFree(buf);
  return rc;
}

// Replaces the node bathymetry (V values) and recomputes the static
// geometry; the state (w, u, v) is kept.
int swe_set_bathymetry(swe_solver* s, const double* z) {
  if (!s || !z) return fail(SWE_ERR_ARG, "null argument");
  if (int rc = set_device(s)) return rc;
  if (int rc = materialize_fluxes(s)) return rc;
  SWE_CHECK(hipMemcpyAsync(const_cast<double*>(s->M.pz), z, sizeof(double) * size_t(s->V),
                           hipMemcpyHostToDevice, s->stream), "hipMemcpy");
  if (s->graph) {
    (void)hipGraphExecDestroy(s->graph);
    s->graph = nullptr;
    s->graph_scheme = -1;
  }
  SWE_CHECK(launch_setup_geometry(s->M, s->G, s->stream), "setup_geometry");
  SWE_CHECK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  return SWE_OK;
}

}  // extern "C"
