/*
 * swe_fvm.h — C ABI of the MI355X (gfx950) explicit finite-volume step for
 * the 2-D shallow-water equations on unstructured triangles.
 *
 * Plain pointers and sizes only.  Every call returns an int status
 * (SWE_OK = 0); on failure swe_last_error() names the cause.  Host arrays are
 * row-major: node xyz is V x 3, cell primitives (w, u, v) are T x 3, edge
 * tables are E x 2, cell tables are T x 3, all indices int64 like the
 * reference's Idx (include/Includes.h:18).
 *
 * A solver handle owns its device buffers and one HIP stream; calls on one
 * handle are not thread-safe.  Different handles are independent.
 *
 * Each entry point cites the reference interface it replaces.
 */
#ifndef SWE_FVM_H
#define SWE_FVM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SWE_OK 0
#define SWE_ERR_ARG 1      /* invalid argument (reference: DomainError)      */
#define SWE_ERR_DEVICE 2   /* HIP runtime failure                             */
#define SWE_ERR_SOLVER 3   /* numerical failure (reference: SolverError)     */
#define SWE_ERR_DOMAIN 4   /* unsupported mesh / boundary tag (DomainError)  */

/* Fluxes::HLL<W> / Fluxes::HLLC<W> (reference include/Fluxes.h:14-15,56-57) */
#define SWE_FLUX_HLL 0
#define SWE_FLUX_HLLC 1
/* Wavespeeds::Rusanov / Davis / Einfeldt (reference include/Fluxes.h:4-10)  */
#define SWE_WAVE_RUSANOV 0
#define SWE_WAVE_DAVIS 1
#define SWE_WAVE_EINFELDT 2
/* Solvers::Euler / SSPRK2 / SSPRK3 (reference include/Solvers.h:6-8)        */
#define SWE_SCHEME_EULER 0
#define SWE_SCHEME_SSPRK2 1
#define SWE_SCHEME_SSPRK3 2

/* Option bits restoring literal HEAD behaviour where the reference is broken */
#define SWE_OPT_HEAD_GRADIENT 1  /* full-wet plane through bathymetry, not w (src/MUSCLObject.cpp:45-64) */
#define SWE_OPT_HEAD_PARTWET2 2  /* keep the transposed, lost writes (src/MUSCLObject.cpp:143,158-159,180) */
/* Linear friction -tau (hu, hv) of the cell average in TimeDisc::RHS; the
 * reference stores tau but never uses it (include/SpaceDisc.h:47,
 * src/TimeDisc.cpp:38-39), so without this bit tau has no effect. */
#define SWE_OPT_FRICTION 16
/* Draining time step with the outflow through each edge, F0 * L, instead of
 * the reference's sum of F0 alone (src/TimeDisc.cpp:56-62), which is a flux
 * per unit length: its A h / sum is not a time, falls below the CFL step
 * in fully wet cells as the mesh is refined and stalls convergence. */
#define SWE_OPT_DRAIN_LENGTH 32
/* Groupings that keep a lake at rest to the bit wherever the two edge states
 * are equal: HLLC's star depth h (a - u) / (a - u*) taken as
 * h ((a - u) / (a - u*)) (include/Fluxes.h:96-108), HLL's flux written
 * Fl + al (ar (Ur - Ul) - (Fr - Fl)) / (ar - al) (include/Fluxes.h:44-52),
 * the RHS adding dt L / A (sgn F - n h_e^2 / 2) per edge instead of the
 * flux and the pressure correction apart (src/TimeDisc.cpp:31,35), and the
 * wall flux taking the inside edge depth h_e, the depth of that correction,
 * instead of the cell mean (src/SpaceDisc.cpp:66-69), and the Shu-Osher
 * stage combinations a0 U0 + a1 U written U + a0 (U0 - U) (a0 + a1 = 1,
 * src/Solvers.cpp:23-25,42-58).  Results
 * differ from the reference's in the last bits.  Not differentiated
 * (swe_adjoint_run and swe_tangent_run refuse it). */
#define SWE_OPT_EXACT_BALANCE 64

/* Boundary tags of edge_elements[e][1] (include/Includes.h:21):
 * SOLID_WALL (src/SpaceDisc.cpp:66-69); FREE_FLOW, declared but
 * unimplemented in the reference: a transmissive edge whose outside state is
 * the inside edge state (the fluxer sees R = L); PERIODIC, declared but
 * never paired there: edges paired by swe_topology_set_periodic.  CUSTOM
 * and an unpaired PERIODIC edge are rejected. */
#define SWE_SOLID_WALL (-1)
#define SWE_FREE_FLOW (-2)
#define SWE_PERIODIC (-3)

typedef struct swe_topology swe_topology;
typedef struct swe_solver swe_solver;

/* Last error message of the calling thread (reference: exception what()). */
const char* swe_last_error(void);

/* ---------------------------------------------------------------- mesh --- */

/* Topology::create / Topology ctor (reference include/TriangMesh.h:16-40,
 * pybind/Topology.cpp:21-27).  Copies (owns) the tables: edge_nodes and
 * edge_elements are E x 2 (edge_elements[e][1] < 0 on the boundary: a tag,
 * of which swe_create accepts SWE_SOLID_WALL, SWE_FREE_FLOW and paired
 * SWE_PERIODIC edges), element_*
 * are T x 3.
 * Indices out of range give SWE_ERR_DOMAIN. */
int swe_topology_create(int64_t num_nodes, int64_t num_edges, const int64_t* edge_nodes,
                        const int64_t* edge_elements, int64_t num_elements,
                        const int64_t* element_nodes, const int64_t* element_edges,
                        const int64_t* element_neighbours, swe_topology** out);

/* Builds the tables of the reference's mesh reader (old TriangMesh(filename),
 * docs/TriangMesh_8h_source.html l.17; conventions fixed by
 * notebooks/topology.dat) from triangles (T x 3, CCW) and boundary lines
 * (B x 2): boundary edges first in line order, the others by first
 * appearance over cells and k = 0..2 of (tp[k], tp[k+1]). */
int swe_topology_build(int64_t num_nodes, int64_t num_triangles, const int64_t* triangle_nodes,
                       int64_t num_boundary_lines, const int64_t* boundary_lines,
                       swe_topology** out);

/* Gmsh 4.1 ASCII reader (old TriangMesh(filename); README.md:21): node xyz
 * (V x 3, tag order) into *xyz (malloc'ed, free with swe_free) and the
 * topology of the 2-D triangles with the 1-D lines as walls. */
int swe_topology_read_msh(const char* path, double** xyz, int64_t* num_nodes,
                          swe_topology** out);

/* Topology::NumNodes / NumEdges / NumTriangles (include/TriangMesh.h:42-44) */
int64_t swe_topology_num_nodes(const swe_topology* t);
int64_t swe_topology_num_edges(const swe_topology* t);
int64_t swe_topology_num_elements(const swe_topology* t);
/* EdgePoints / EdgeTriangs / TriangPoints / TriangEdges / TriangTriangs
 * (include/TriangMesh.h:46-52), whole tables; any pointer may be NULL. */
/* PERIODIC boundaries (include/Includes.h:21, declared there and never
 * paired): boundary edges pairs[2k] and pairs[2k+1] are tagged PERIODIC
 * (-3) and paired; the outside state of each is its partner's inside slot,
 * the flux of a pair is evaluated once (from the pair's lower edge) and
 * applied with opposite signs, and the partner cell is the neighbour of
 * the draining time step.  Pairs should join edges of equal length and
 * opposite outward normals (a translation of the mesh boundary). */
int swe_topology_set_periodic(swe_topology* t, int64_t n, const int64_t* pairs);
/* partner[e] = the edge paired with edge e, -1 for unpaired edges. */
int swe_topology_periodic(const swe_topology* t, int64_t* partner);
int swe_topology_copy(const swe_topology* t, int64_t* edge_nodes, int64_t* edge_elements,
                      int64_t* element_nodes, int64_t* element_edges,
                      int64_t* element_neighbours);
void swe_topology_destroy(swe_topology* t);
void swe_free(void* p);

/* StructTriangMesh(ni, nj, h) (reference include/StructTriangMesh.h:4-15):
 * ni x nj squares of side h from (x0, y0), four CCW triangles per square
 * (bottom, right, top, left) around a centre node.  Grid nodes first
 * (row-major, (ni+1)(nj+1)), then centre nodes; walls are the outer edges. */
int swe_mesh_struct(int64_t ni, int64_t nj, double h, double x0, double y0, double** xyz,
                    int64_t* num_nodes, swe_topology** out);

/* Seeded Delaunay triangulation of an n x n jittered grid on
 * [x0, x0+L] x [y0, y0+L] (interior nodes moved by U(-j, j) * L / n), cells
 * reordered along a Morton curve of their centroids and nodes renumbered by
 * first appearance; the tables are then built as swe_topology_build does. */
int swe_mesh_jittered(int64_t n, double jitter, uint64_t seed, double x0, double y0, double L,
                      double** xyz, int64_t* num_nodes, swe_topology** out);

/* -------------------------------------------------------------- solver --- */

/* SpaceDisc(fluxer, Domain(geometry, topology), v0, cor, tau)
 * (reference include/SpaceDisc.h:24, include/Bathymetry.h:14).  xyz is the
 * node geometry (x, y, bathymetry z), V x 3.  device: HIP ordinal.  The
 * tables are validated before any device work (SWE_ERR_ARG for indices out
 * of range or a mesh without triangles, SWE_ERR_DOMAIN for unsupported
 * boundary tags). */
int swe_create(const swe_topology* topo, const double* xyz, int flux, int wavespeed, double cor,
               double tau, int options, int device, swe_solver** out);
void swe_destroy(swe_solver* s);

/* VolumeField write / read (include/ValueField.h:33-44): primitive (w, u, v),
 * T x 3 host arrays.  Writes store the values as given (no assigner). */
int swe_set_state(swe_solver* s, const double* prim);
int swe_get_state(swe_solver* s, double* prim);

/* SpaceDisc::ComputeInterfaceValues (src/SpaceDisc.cpp:33-52). */
int swe_compute_interface_values(swe_solver* s);
/* SpaceDisc::ComputeFluxes (src/SpaceDisc.cpp:54-74); *r receives
 * GetMinLenToWavespeed() (include/SpaceDisc.h:31) when non-NULL. */
int swe_compute_fluxes(swe_solver* s, double* r);

/* Getters of SpaceDisc (include/SpaceDisc.h:26-31), downloaded to host:
 * edge slots (w, u, v) 2E x 3 (slot 2e + (from < to), include/ValueField.h:70-75),
 * source slots 2E x 3 (column 0 unused), fluxes E x 3, node maximum V. */
int swe_get_edge_field(swe_solver* s, double* edg);
int swe_get_src_field(swe_solver* s, double* src);
int swe_get_fluxes(swe_solver* s, double* f);
int swe_get_max_wp(swe_solver* s, double* max_wp);
/* MUSCLObject::IsDryCell / IsFullWetCell (src/MUSCLObject.cpp:13-29) for
 * every cell: 0 dry, 1 partially wet, 2 fully wet. */
int swe_get_cell_class(swe_solver* s, int8_t* cls);

/* TimeDisc::RHS(i, dt) (src/TimeDisc.cpp:3-41) after ComputeFluxes. */
int swe_rhs(swe_solver* s, int64_t cell, double dt, double* out3);
/* MUSCLObject::ReconstructDryCell / PartWetCell1 / FullWetCell /
 * PartWetCell2 (include/MUSCLObject.h:57-60, src/MUSCLObject.cpp:31-191) of
 * one cell of the current state (kind 0, 1, 2, 3), evaluated on the device
 * with the step's own code after recomputing the interface values (pass 2
 * reads the pass-1 node maximum).  out5 = origin (w, u, v) at the centroid
 * and the gradient of w (x, y); u and v have zero gradient in every
 * reconstruction.  A wall neighbour of a full-wet plane is the cell itself
 * (the reference reads index -1 there). */
int swe_reconstruct(swe_solver* s, int64_t cell, int kind, double* out5);
/* Fluxes::HLL<W> / HLLC<W> (include/Fluxes.h:14-111) of edge e seen from
 * cell `from` towards cell `to` (to < 0: a boundary edge, the right state
 * is the left one), flux SWE_FLUX_* and wave-speed estimate SWE_WAVE_*, with
 * the edge slots of the last interface values: out3 = flux in the normal
 * outward of `from`; *r (may be null) = min(*r, min(2 A_from / L, 2 A_to /
 * L) / wave speed), left unchanged when the flux returns before the
 * speeds (dry edge, vanishing speeds), as in the reference. */
int swe_edge_flux(swe_solver* s, int flux, int wavespeed, int64_t e, int64_t from, int64_t to, double* r,
                  double* out3);

/* Solvers::Euler / SSPRK2 / SSPRK3 (src/Solvers.cpp:5-59) repeated nsteps
 * times on the device without host synchronisation.  dt > 0: fixed step;
 * dt <= 0: CFL mode, each step uses cfl * r of its first flux pass.
 * Asynchronous: call swe_synchronize (or swe_get_*) before reading. */
int swe_run(swe_solver* s, int scheme, double dt, double cfl, int64_t nsteps);
/* dt of each step of the last swe_run (nsteps doubles) and simulated time. */
int swe_get_dts(swe_solver* s, double* dts, int64_t nsteps);
int swe_synchronize(swe_solver* s);
/* TimeDisc::CFLdt (include/TimeDisc.h:13): 0.15 * last r. */
int swe_cfl_dt(swe_solver* s, double* dt);
/* SpaceDisc::GetMinLenToWavespeed (include/SpaceDisc.h:31): r of the last
 * flux pass (of swe_compute_fluxes or of the last stage of swe_run). */
int swe_min_len_to_wavespeed(swe_solver* s, double* r);

/* Sizes and raw device pointers (state w, u, v SoA) for zero-copy interop;
 * swe_device_state refuses a solver whose state layout is not SoA
 * (swe_set_layout). */
int64_t swe_num_cells(const swe_solver* s);
int64_t swe_num_edges(const swe_solver* s);
int64_t swe_num_nodes(const swe_solver* s);
int swe_device_state(swe_solver* s, double** w, double** u, double** v);

/* ------------------------------------------------------------ ensembles --- */

/* Ensemble of n independent meshes advanced together (no reference
 * counterpart: the reference advances one SpaceDisc at a time).  The
 * members are concatenated; state arrays (swe_set_state / swe_get_state)
 * hold the members one after another (swe_batch_info gives the offsets).
 * Each member has its own CFL step dt_k = cfl * r_k of its own first flux
 * pass, clipped to land on t_end, and stops there. */
int swe_batch_create(int64_t n, const swe_topology* const* topos, const double* const* xyz,
                     int flux, int wavespeed, double cor, double tau, int options, int device,
                     swe_solver** out);
/* Sets every member's time to t0 and makes it active again. */
int swe_batch_reset(swe_solver* s, double t0);
/* Steps until every member reached t_end or max_steps steps were taken
 * (checked every 64 steps).  use_graph: replay one captured HIP graph per
 * step instead of launching the kernels one by one. */
int swe_batch_run(swe_solver* s, int scheme, double cfl, double t_end, int64_t max_steps,
                  int use_graph, int64_t* steps_done, int64_t* n_active);
/* Member count and cell / edge / node offsets (n + 1 entries each). */
int swe_batch_info(swe_solver* s, int64_t* n, int64_t* cell_off, int64_t* edge_off,
                   int64_t* node_off);
/* Each member's simulated time and number of steps taken. */
int swe_batch_time(swe_solver* s, double* t, int64_t* steps);

/* Time, with HIP events on the solver's stream, nsteps steps of swe_run:
 * *ms = total milliseconds; kernel_ms (optional, 7 doubles) = average
 * per-launch duration of the reconstruction, wet/dry, flux, draining,
 * update, fused flux-draining-update and fused draining-update (after the
 * flux pass of a CFL step's first stage) kernels (0 for a kernel the run did
 * not use), measured by bracketing each launch with events. */
int swe_time_run(swe_solver* s, int scheme, double dt, double cfl, int64_t nsteps, double* ms,
                 double* kernel_ms);

/* ----------------------------------------------------------- diagnostics --- */

/* ComputeIntegrals (src/SpaceDisc.cpp:77-103, commented out in the
 * reference) of the current state: out3 = (mass, kinetic energy, potential
 * energy) = sums over cells of area * (h, 1/2 h |u|^2, 1/2 h^2 + h b), with
 * partially wet cells integrating the energies of their reconstruction by
 * TriangAverage<2,10>.  Recomputes the interface values first.  The sum is
 * a fixed-order tree (identical run to run); per_cell (optional, 3 T
 * doubles) receives the cell terms.  Single-mesh handles only. */
int swe_integrals(swe_solver* s, double* out3, double* per_cell);
/* Switches the tile kernel off (on = 0) or back on (no reference
 * counterpart).  When on (the default, on meshes whose tiles fit the
 * kernel), every stage whose step size is known before its flux pass (fixed
 * dt; stages 2 and 3 of a CFL step) runs as reconstruction + one fused
 * kernel instead of reconstruction + flux + draining + update, and the first
 * stage of a CFL step as reconstruction + flux + one kernel for the
 * draining and the update; results are identical bit for bit.  *active (may
 * be null) receives whether the kernel will be used. */
int swe_set_fused(swe_solver* s, int on, int* active);
/* Allocation-free swe_integrals.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same results as
 * swe_integrals, bit for bit, but out3_dev (3 doubles) and per_cell_dev
 * (3 T doubles, may be null) are DEVICE pointers written asynchronously on
 * the solver's stream (swe_synchronize before reading them on the host),
 * and the block partials live in a static per-GPU scratch array of
 * SWE_NOALLOC_MAX_CELLS / 256 x 3 doubles (1.5 MiB, held until the process
 * ends): meshes above SWE_NOALLOC_MAX_CELLS cells are refused with
 * SWE_ERR_ARG.  Calls no hipMalloc / hipFree.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams or
 * threads) race.  Single-mesh handles only. */
#define SWE_NOALLOC_MAX_CELLS 16777216
int swe_integrals_noalloc(swe_solver* s, double* out3_dev, double* per_cell_dev);
/* Final per-cell reconstruction after swe_compute_interface_values: level
 * w0 at the centroid and gradient (gx, gy); velocity is the cell's (zero
 * for dry cells).  Any pointer may be null. */
int swe_get_muscl(swe_solver* s, double* w0, double* gx, double* gy);

/* Failure detection (no reference counterpart: the reference has no NaN or
 * depth checks): out5 = (number of cells with a non-finite w, u or v,
 * min h, max h, max |u|, max |u| + sqrt(h)), the speeds over wet cells.
 * Exact (min / max / counts) and identical run to run. */
int swe_state_stats(swe_solver* s, double* out5);

/* ------------------------------------------------------------- adjoint --- */

/* Adjoint of nsteps steps (no reference counterpart).  From the current
 * state U_0, runs the steps with fixed sizes (dt > 0 for every step, or
 * dts[nsteps], e.g. the sizes a CFL run logged, then held fixed), storing
 * U_n every checkpoint_every steps (<= 0 or >= nsteps: one segment); U_N is
 * the current state on return.  Then propagates final_bar (T x 3 row-major
 * adjoint of U_N's w, u, v) back through every stage into state0_bar (T x
 * 3, adjoint of U_0) and z_bar (V, adjoint of the node bathymetry; may be
 * NULL).  Each segment, last first, is recomputed from its checkpoint with
 * every stage's intermediates kept on the device (about 81 bytes per cell
 * and stage plus 24 per cell and stage state), then reversed; the last
 * segment's recomputation is the forward run.  Branch decisions (wet/dry,
 * limiter, wave-speed selections) are the forward ones; the Riemann
 * solver is differentiated in reverse mode (the transpose of its
 * dual-number derivative) and the partially wet reconstructions by local
 * Jacobians on dual numbers (implicit derivatives of the cube-root and
 * bisection levels); the node maxima's adjoints go to the cells attaining
 * them.  The
 * HEAD-compatibility options give SWE_ERR_ARG.  ms (optional, 2 doubles):
 * forward sweep and reverse sweep (recomputation included) milliseconds.
 * Single-mesh handles only. */
int swe_adjoint_run(swe_solver* s, int scheme, double dt, const double* dts, int64_t nsteps,
                    int64_t checkpoint_every, const double* final_bar, double* state0_bar,
                    double* z_bar, double* ms);
/* Timing of the last swe_adjoint_run on this handle (no reference
 * counterpart): out3 = milliseconds of the forward sweep, of the reverse
 * sweep's recomputation of its segments from their checkpoints (forward
 * work that checkpointing costs), and of the rest of the reverse sweep
 * (the reverse stages; the two last add up to swe_adjoint_run's ms[1]);
 * zeros before the first run. */
int swe_adjoint_timing(const swe_solver* s, double* out3);
/* Tangent-linear run (no reference counterpart): the directional
 * derivative of nsteps steps of fixed sizes (dt > 0, or dts[nsteps]) along
 * state0_dot (T x 3 row-major tangent of the initial w, u, v) and z_dot (V
 * tangent of the node bathymetry; NULL: zero), computed by the stage kernels
 * built on dual numbers.  Their value parts are the forward step's, bit for
 * bit: U_N is the current state on return.  final_dot (T x 3) receives the
 * tangent of U_N.  With the adjoint it satisfies <final_bar, final_dot> =
 * <state0_bar, state0_dot> + <z_bar, z_dot>.  ms (optional, 1 double):
 * milliseconds.  Single-mesh handles only. */
int swe_tangent_run(swe_solver* s, int scheme, double dt, const double* dts, int64_t nsteps,
                    const double* state0_dot, const double* z_dot, double* final_dot, double* ms);
/* Replaces the node bathymetry z (V values) and recomputes the static
 * geometry; the state is kept (Domain's z, include/Bathymetry.h:43). */
int swe_set_bathymetry(swe_solver* s, const double* z);

/* ---------------------------------------------------------------- tuning --- */

/* Work items per thread of the flux kernel (1, 2 or 4 edges) and of the
 * update kernel (1 or 2 cells); no reference counterpart (the reference is
 * serial).  Results do not depend on the choice, only the speed does.
 * Defaults: 2 and 1. */
int swe_set_tuning(swe_solver* s, int flux_items, int update_items);
int swe_get_tuning(const swe_solver* s, int* flux_items, int* update_items);
/* Threads per workgroup of the flux and update kernels (256 or 512); no
 * reference counterpart.  Results do not depend on the choice.  Defaults:
 * 256 and 256. */
int swe_set_tuning_blocks(swe_solver* s, int flux_block, int update_block);
int swe_get_tuning_blocks(const swe_solver* s, int* flux_block, int* update_block);
/* Times every instantiation (work items x block size) reps times on the
 * current state, keeps the fastest and returns its work items (the block
 * sizes through swe_get_tuning_blocks); times (optional, 5 doubles)
 * receives the average per-launch milliseconds of flux x1, x2, x4 and
 * update x1, x2, each at its faster block size.  The state is unchanged;
 * the reconstruction and flux fields afterwards are those of the current
 * state. */
int swe_autotune(swe_solver* s, int reps, int* flux_items, int* update_items, double* times);

/* Layout of the cell state (w, u, v) in device memory, a dimension of the
 * autotuner: block 0 is SoA (three arrays); a power of two B up to 64 stores
 * blocks of B cells as B values of w, then of u, then of v (AoSoA; B = 1 is
 * AoS).  Results do not depend on it.  Changing it rewrites the current
 * state on the device and drops a captured ensemble graph.  Pointers from
 * swe_device_state, and swe_adjoint_run / swe_tangent_run, need SoA.  No
 * counterpart in the reference (its VolumeField is an Eigen 3 x T array,
 * include/ValueField.h:17-44). */
int swe_set_layout(swe_solver* s, int block);
int swe_get_layout(const swe_solver* s, int* block);
/* Tunes the layout with the work-item variants: for each candidate layout
 * (SoA, 64, 8, 1) swe_autotune, then the first stage of a CFL step and, for
 * a solver that fuses its later stages, one fused stage timed on the live
 * state;
 * keeps the layout and variants of least cost.  cost_ms (may be null)
 * receives SWE_LAYOUT_CANDIDATES costs in that order. */
#define SWE_LAYOUT_CANDIDATES 4
int swe_autotune_layout(swe_solver* s, int reps, int* block, double* cost_ms);

#ifdef __cplusplus
}
#endif

#endif /* SWE_FVM_H */
