"""Throughput of the explicit finite-volume step on one MI355X.

Workload (default): wet/dry circular dam break on a 16M-triangle seeded
Delaunay mesh of [0, 1]^2 (n = 2828 jittered grid, Morton-ordered), random
Fourier bathymetry (64 modes, amplitude 0.1), h = 1 inside r < 0.2, h = 0.1
outside, 5 % dry pockets; FP64, HLLC<Einfeldt>, SSP-RK2, CFL mode 0.15.

One step = one SSP-RK2 time step of every cell; value = cell-updates/s with
the state resident in HBM.  Prints one JSON line.  The same line carries
the batched-ensemble and adjoint/forward ratios of shortened runs of the
256-member ensemble and of the 32-member bathymetry inversion
(`ensemble`, `adjoint`), the roofline of the dominant kernel and of the
whole step against the compulsory bytes of a cell-update, and CPU rows
(the oracle, one thread, on the host of the GPU) for the Thacker and
lake-at-rest configurations.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 2828]
    python bench.py --ensemble | --adjoint      (the full runs, one line each)
"""
import argparse
import gc
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
# Compulsory bytes per cell of one SSP-RK2 stage: read the stage state
# (24), write the new state (24), 3 neighbour + 3 node ids (24), node
# records at V/T ~ 0.5 (12); stage 2 also reads U^n (24).
STAGE_BYTES = (84.0, 108.0)
STEP_BYTES_PER_CELL = sum(STAGE_BYTES)  # 192


def host_cpu():
    """CPU model and the cores this process may use (lscpu's model name)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def layout_name(block):
    """The state layout of swe_set_layout: 0 = SoA, B = AoSoA blocks of B cells."""
    return "soa" if block == 0 else f"aosoa{block}"


def kernel_own_bytes(T, E, V):
    """Bytes of every array a kernel reads or writes, once each (int32
    indices, FP64 values), per launch: the kernel's own implementation
    arrays, intermediates included (not the compulsory bytes of a
    cell-update, which the roofline uses)."""
    nodes = 32.0 * V  # the interleaved node records
    return {
        # w, 32-byte record, vertex ids + flags, neighbour ids, nodes in;
        # MUSCL (w0, gx, gy) + class out
        "recon1": T * (8 + 32 + 16 + 12) + nodes + T * (24 + 1),
        # class of every cell (pass 2 itself is sparse)
        "recon2": T * 1.0,
        # edge ends/cells, CFL length, per-cell MUSCL + class + u, v +
        # centroid, nodes in; F out
        "flux": E * (16 + 8) + T * (24 + 1 + 16 + 16) + nodes + E * 24,
        # w, tp, te, cflag, F0, nodes in; dtd out
        "drain": T * (8 + 12 + 12 + 4) + E * 8 + nodes + T * 8,
        # state (+ U0), tp, te, tt, cflag, MUSCL, class, F, dtd, nodes in; state out
        "update": T * (24 + 24 + 36 + 4 + 24 + 1 + 8) + E * 24 + nodes + T * 24,
        # state (+ U0), vertex ids + flags, tile codes, MUSCL, class,
        # centroids, edge records, nodes in; state out (fluxes and draining
        # steps stay in LDS)
        "fused": T * (24 + 24 + 16 + 12 + 24 + 1 + 16) + E * 16 + nodes + T * 24,
        # as the fused kernel without the per-edge MUSCL / centroid gathers:
        # each edge's flux (24 B) and evaluation index in
        "fused_ff": T * (24 + 24 + 16 + 12 + 24 + 1) + E * (4 + 24) + nodes + T * 24,
    }




def load_traffic(workload):
    """HBM bytes per launch of each kernel and launches per step, from the
    committed rocprofv3 --pmc summary (profiles/pmc_traffic.json), if any."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        data = json.load(f)
    return data.get(workload, {})


def cpu_baseline(dom, prim, flux, wave, reps=3):
    """The oracle (scalar C++ restatement, 1 thread) on the same state: one
    SSP-RK2 step in CFL mode after one warm-up step, median of reps."""
    from oracle import oracle as O
    orc = O.OracleSolver(dom.geometry, dom.topology.tables(), flux=flux, wave=wave)
    orc.set_state(prim)
    orc.time_run("SSPRK2", 1, dt=0.0, cfl=0.15)  # warm-up
    times = []
    for _ in range(reps):
        orc.set_state(prim)
        times.append(orc.time_run("SSPRK2", 1, dt=0.0, cfl=0.15))
    el = statistics.median(times)
    T = prim.shape[0]
    cpu = host_cpu()
    return {"value": T / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "host_cpu": cpu["model"], "host_logical_cpus": cpu["logical_cpus"],
            "sample": f"1 SSP-RK2 step (CFL mode) of the same {T}-cell state after a warm-up step, "
                      f"median of {reps} ({', '.join(f'{t:.2f}' for t in times)} s), oracle C++ restatement "
                      f"-O3 -ffp-contract=off, 1 thread of {cpu['model']}"}


def cpu_rows(fl):
    """The CPU-against-GPU rows: the oracle (1 thread) against the GPU on the
    Thacker configuration (40,000 cells, one period, CFL 0.5, in full) and
    on the 1M-cell lake at rest (3 steps sampled, HLLC<Einfeldt>)."""
    import swe_fvm_amd as swe
    from oracle import oracle as O
    from swe_fvm_amd import workloads as W
    from swe_fvm_amd.ensemble import Ensemble
    rows = []
    # Thacker, 40,000 cells: the GPU runs to the period with per-step CFL dt (the last step
    # clipped); the oracle replays the same steps
    dom, prim, th = W.thacker(n=100, quad=100)
    T = prim.shape[0]
    ens = Ensemble(fl, [dom], [prim])
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run("SSPRK2", cfl=0.5, t_end=th.period, graph=True)
    ens.synchronize()
    g_s = time.perf_counter() - t0
    _, steps = ens.times()
    n = int(steps[0])
    orc = O.OracleSolver(dom.geometry, dom.topology.tables(), "HLLC", "Einfeldt")
    orc.set_state(prim)
    t0 = time.perf_counter()
    tk = 0.0
    for _ in range(n):
        orc.compute_interface_values()
        d = 0.5 * orc.compute_fluxes()
        if d < th.period - tk:
            orc.run("SSPRK2", 1, dt=0.0, cfl=0.5)
            tk += d
        else:
            orc.run("SSPRK2", 1, dt=th.period - tk)
            tk = th.period
    c_s = time.perf_counter() - t0  # includes the oracle's extra r pass per step
    same = bool(np.array_equal(ens.states()[0], orc.state()))
    rows.append({"config": "thacker_100x100_T40000", "cells": T, "steps": n, "full": True,
                 "cpu_cell_updates_per_s": T * n / c_s, "gpu_cell_updates_per_s": T * n / g_s,
                 "gpu_over_cpu": c_s / g_s, "bitwise_equal": same,
                 "note": "GPU: one graph-captured ensemble step per step (launch-bound at 40k cells); CPU time "
                         "includes one extra interface-value and flux pass per step for the clipped-step replay"})
    del ens
    # lake at rest, 999,698 cells: 3 steps sampled on the CPU, 100 on the GPU
    dom, prim = W.lake_at_rest(n=707, seed=1)
    T = prim.shape[0]
    orc = O.OracleSolver(dom.geometry, dom.topology.tables(), "HLLC", "Einfeldt")
    orc.set_state(prim)
    orc.time_run("SSPRK2", 1, dt=0.0, cfl=0.15)
    c_s = orc.time_run("SSPRK2", 3, dt=0.0, cfl=0.15)
    sd = swe.SpaceDisc(fl, dom, prim)
    sd.run("SSPRK2", 5, dt=0.0, cfl=0.15)
    g_ms = sd.time_run("SSPRK2", 100, dt=0.0, cfl=0.15)
    rows.append({"config": "lake_at_rest_T999698", "cells": T, "steps_cpu": 3, "steps_gpu": 100, "full": False,
                 "cpu_cell_updates_per_s": T * 3 / c_s, "gpu_cell_updates_per_s": T * 100 / (g_ms * 1e-3),
                 "gpu_over_cpu": (T * 100 / (g_ms * 1e-3)) / (T * 3 / c_s)})
    del sd
    return rows


def ensemble_measure(n_members, cells, t_end, seq_steps, host_stepped=True, tune_reps=3, chunk=0):
    """n_members seeded members (log-uniform sizes, ~cells in total),
    HLLC<Einfeldt>, SSP-RK2, per-member CFL 0.15 steps to t_end: the batched
    ensemble (one graph-captured step for all active members) against the
    same members advanced one after the other with the single-mesh solver
    (device-stepped: seq_steps steps per call; host-stepped: one
    Solvers.SSPRK2 call per step, as the reference's drivers step).  With
    chunk > 0 the batched run goes chunk steps per call and reports each
    chunk's active members and rate (where the run loses throughput as
    members finish)."""
    import swe_fvm_amd as swe
    from swe_fvm_amd import workloads as W
    from swe_fvm_amd.ensemble import Ensemble
    t0 = time.time()
    mem = W.ensemble(n_inst=n_members, total=cells, seed=11)
    T = np.array([p.shape[0] for _, p in mem])
    print(f"[bench] ensemble {len(mem)} members, {T.sum()} cells ({T.min()}..{T.max()}) in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    ens = Ensemble(fl, [d for d, _ in mem], [p for _, p in mem])
    tune = ens.autotune_layout(reps=tune_reps)  # work items, workgroup size and state layout
    ens.run("SSPRK2", cfl=0.15, t_end=t_end, max_steps=2, graph=True)  # warm-up + capture
    ens.reset(0.0)
    ens.set_states([p for _, p in mem])
    ens.synchronize()
    chunks = []
    t1 = time.perf_counter()
    if chunk > 0:
        done, act, prev = 0, len(mem), np.zeros(len(mem), np.int64)
        while act > 0:
            tc = time.perf_counter()
            d, act = ens.run("SSPRK2", cfl=0.15, t_end=t_end, max_steps=chunk, graph=True)
            ens.synchronize()
            wc = time.perf_counter() - tc
            st = ens.times()[1]
            upd = float(np.sum((st - prev) * T))
            prev = st.copy()
            done += d
            chunks.append({"steps": done, "active_after": act, "cell_updates": upd, "seconds": wc,
                           "rate": upd / wc})
            print(f"[bench] chunk to step {done}: {act} active, {upd / wc:.3e}/s", file=sys.stderr, flush=True)
            if d == 0:
                break
    else:
        done, act = ens.run("SSPRK2", cfl=0.15, t_end=t_end, graph=True)
    ens.synchronize()
    wall_b = time.perf_counter() - t1
    tt, steps = ens.times()
    updates = float(np.sum(T * steps))
    batched = updates / wall_b
    print(f"[bench] batched: {done} steps, {updates:.3e} cell-updates in {wall_b:.2f}s "
          f"-> {batched:.3e}/s, all at t_end: {bool(np.all(tt == t_end))}", file=sys.stderr, flush=True)
    del ens
    gc.collect()
    seq_updates, seq_s = 0.0, 0.0
    for dom, prim in mem:
        sd = swe.SpaceDisc(fl, dom, prim)
        sd.run("SSPRK2", 1, dt=0.0, cfl=0.15)
        sd.synchronize()
        t2 = time.perf_counter()
        sd.run("SSPRK2", seq_steps, dt=0.0, cfl=0.15)
        seq_s += time.perf_counter() - t2
        seq_updates += prim.shape[0] * seq_steps
        del sd
    sequential = seq_updates / seq_s
    print(f"[bench] sequential: {seq_updates:.3e} cell-updates in {seq_s:.2f}s -> {sequential:.3e}/s",
          file=sys.stderr, flush=True)
    out = {"value": batched / sequential, "unit": "ratio (batched / sequential device-stepped cell-updates/s)",
           "batched_cell_updates_per_s": batched, "sequential_cell_updates_per_s": sequential,
           "batched": {"steps": done, "members_active_at_end": act, "cell_updates": updates, "seconds": wall_b,
                       "t_end": t_end, "all_at_t_end": bool(np.all(tt == t_end)), "graph": True},
           "sequential": {"steps_per_member": seq_steps, "cell_updates": seq_updates, "seconds": seq_s},
           "chunks": chunks or None,
           "config": {"workload": f"ensemble_{len(mem)}_members_T{int(T.sum())}", "members": len(mem),
                      "cells_min": int(T.min()), "cells_max": int(T.max()), "scheme": "SSPRK2",
                      "flux": "HLLC<Einfeldt>", "dt": "per-member CFL 0.15",
                      "tuning": {"flux_edges_per_thread": tune["flux_items"],
                                 "update_cells_per_thread": tune["update_items"],
                                 "flux_block_threads": tune["flux_block"],
                                 "update_block_threads": tune["update_block"],
                                 "state_layout": layout_name(tune["layout"]),
                                 "stage_ms_per_layout": tune["stage_ms"]}}}
    if host_stepped:
        hs_updates, hs_s = 0.0, 0.0
        for dom, prim in mem:
            sd = swe.SpaceDisc(fl, dom, prim)
            td = swe.TimeDisc(sd)
            sd.ComputeInterfaceValues()
            sd.ComputeFluxes()
            dt = td.CFLdt()
            swe.Solvers.SSPRK2(td, dt)
            t2 = time.perf_counter()
            for _ in range(seq_steps):
                swe.Solvers.SSPRK2(td, dt)
            hs_s += time.perf_counter() - t2
            hs_updates += prim.shape[0] * seq_steps
            del sd
        out["sequential_host_stepped_cell_updates_per_s"] = hs_updates / hs_s
        out["ratio_vs_host_stepped"] = batched / (hs_updates / hs_s)
    return out


def bench_ensemble(args):
    """The full 256-member run (t_end = 1.0) as one JSON line."""
    r = ensemble_measure(args.members, args.ensemble_cells, args.t_end, args.seq_steps, chunk=args.chunk)
    r.update({"metric": "batched/sequential aggregate FP64 cell-updates/s", "higher_is_better": True,
              "dtype": "f64", "data": "synthetic", "n_gpus": 1})
    print(json.dumps(r))


def adjoint_measure(B=32, n=354, N=2000, dt=2e-4, checkpoint=50, nfd=16, check_steps=100, fused=True):
    """Bathymetry inversion, a batch of B members on the n x n-grid mesh
    (n = 354: 250,632 cells) stacked into one domain (every kernel advances
    all members in one launch), HLLC<Einfeldt>, SSP-RK2, fixed dt, N steps,
    checkpoints every `checkpoint` steps.  The gradient of the summed gauge
    misfits with respect to the initial state and the node bathymetry comes
    from the torch.autograd.Function (ShallowWaterRun).  Reports the plain
    batched forward, the adjoint's forward sweep (checkpoints) and reverse
    sweep (segment recomputation + reverse stages); value = reverse sweep /
    forward, the cost of the adjoint step per forward step, recomputation
    included; gradient_cost_ratio = (forward sweep + reverse sweep) /
    forward.  Checks: member 0 run alone reproduces its slice of the batch
    bit for bit; in nfd random directions (bathymetry and initial surface
    together) its gradient satisfies the duality with the tangent of the
    dual-number build of the forward over the full horizon, and is compared
    with central differences over the full horizon and over check_steps
    steps, at eps and eps/4 (a pair that disagrees crossed one of the
    step's switches, the limiter's all-or-nothing decisions:
    inconclusive)."""
    import torch

    import swe_fvm_amd as swe
    from swe_fvm_amd import adjoint as AD
    from swe_fvm_amd import workloads as W
    t0 = time.time()
    ck = checkpoint if checkpoint > 0 else N
    d = W.adjoint_inversion_batch(members=B, n=n)
    dom1, prim1 = d["member"]
    T1, V1 = prim1.shape[0], dom1.geometry.shape[0]
    big, off = swe.stack_domains([dom1] * B)
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    prim = np.tile(prim1, (B, 1))
    z0 = np.tile(d["z"], B)
    sd = swe.SpaceDisc(fl, big, prim)
    fused = sd.set_fused(fused)  # the forward's stages and the reverse sweep's recomputation
    # targets: one batched run over the members' true bathymetries
    sd.set_bathymetry(np.concatenate(d["z_true"]))
    sd.SetVolField(prim)
    sd.run("SSPRK2", N, dt=dt)
    fin = sd.GetVolField()
    members = [(g + off["cells"][k], fin[g + off["cells"][k], 0].copy()) for k, g in enumerate(d["gauges"])]
    print(f"[bench] adjoint batch {B} x {T1} cells, {N} steps, setup {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    sd.set_bathymetry(z0)
    sd.SetVolField(prim)
    ms_fwd = sd.time_run("SSPRK2", N, dt=dt)
    fin_fwd = sd.GetVolField()
    p = torch.tensor(prim, requires_grad=True)
    zt = torch.tensor(z0, requires_grad=True)
    out = AD.ShallowWaterRun.apply(p, zt, sd, N, "SSPRK2", dt, ck)
    gidx = torch.from_numpy(np.concatenate([m[0] for m in members]))
    tgt = torch.from_numpy(np.concatenate([m[1] for m in members]))
    loss = ((out[gidx, 0] - tgt) ** 2).sum()
    loss.backward()
    ms = sd.last_adjoint_ms
    split = sd.adjoint_timing()  # the reverse sweep's recomputation and reverse stages
    gp, gz = p.grad.numpy(), zt.grad.numpy()
    print(f"[bench] forward {ms_fwd:.1f} ms, adjoint sweeps {ms[0]:.1f} + {ms[1]:.1f} ms "
          f"(recomputation {split['recompute_ms']:.1f}, reverse stages {split['reverse_stages_ms']:.1f})",
          file=sys.stderr, flush=True)
    # member 0 alone: its slice of the batch, bit for bit
    sd1 = swe.SpaceDisc(fl, dom1, prim1)
    sd1.set_fused(fused)
    z1 = d["z"]

    def run1(zz, p0, nsteps):
        sd1.set_bathymetry(zz)
        sd1.SetVolField(p0)
        sd1.run("SSPRK2", nsteps, dt=dt)
        return sd1.GetVolField()
    slice_same = bool(np.array_equal(run1(z1, prim1, N), fin_fwd[:T1]))
    g0, t0v = d["gauges"][0], members[0][1]
    taper = W.wall_taper(dom1.geometry[:, 0]) * W.wall_taper(dom1.geometry[:, 1])
    rng = np.random.default_rng(3)
    c1 = dom1.centroids()
    dirs = []
    for _ in range(nfd):  # smooth directions: bathymetry and initial surface
        dp = np.zeros_like(prim1)
        dp[:, 0] = W.smooth_field(c1[:, 0], c1[:, 1], rng, 1e-2)
        dirs.append((W.smooth_field(dom1.geometry[:, 0], dom1.geometry[:, 1], rng, 1e-2) * taper, dp))

    def fd_checks(nsteps, fnc, adz, adp):
        res = []
        for dz, dp in dirs:
            J = lambda e: fnc(run1(z1 + e * dz, prim1 + e * dp, nsteps))[0]
            cd = lambda e: (J(e) - J(-e)) / (2 * e)
            a, b = cd(1e-6), cd(2.5e-7)
            ad = float(np.sum(adz * dz) + np.sum(adp * dp))
            ok = abs(a - b) <= 1e-4 * abs(a)
            res.append({"adjoint": ad, "fd": a, "fd_quarter_eps": b, "conclusive": bool(ok),
                        "rel": abs(ad - a) / abs(a) if ok and a != 0 else None})
        rels = [r["rel"] for r in res if r["rel"] is not None]
        return {"steps": nsteps, "directions": len(res), "conclusive": len(rels),
                "max_rel": max(rels) if rels else None, "results": res}
    f0 = AD.gauge_misfit(g0, t0v)
    # duality with the dual-number build of the forward, full horizon:
    # <dJ/dU_N, U_N'> = <dJ/dU_0, U_0'> + <dJ/dz, z'> for every direction
    ybar = f0(fin_fwd[:T1])[1]
    dual = []
    tl_ms = 0.0
    for dz, dp in dirs:
        sd1.set_bathymetry(z1)
        sd1.SetVolField(prim1)
        tdot, tms = sd1.tangent_run(dp, N, z_dot=dz, scheme="SSPRK2", dt=dt)
        tl_ms += tms
        lhs = float(np.sum(ybar * tdot))
        rhs = float(np.sum(gp[:T1] * dp) + np.sum(gz[:V1] * dz))
        dual.append({"tangent": lhs, "adjoint": rhs, "rel": abs(lhs - rhs) / max(abs(lhs), 1e-300)})
    tangent = {"steps": N, "directions": len(dual), "max_rel": max((r["rel"] for r in dual), default=None),
               "tangent_run_ms": tl_ms / max(len(dual), 1), "results": dual}
    fd_full = fd_checks(N, f0, gz[:V1], gp[:T1])
    nc = check_steps
    fc = AD.gauge_misfit(g0, run1(d["z_true"][0], prim1, nc)[g0, 0].copy())
    _, s0c, zbc, _ = AD.run_and_adjoint(sd1, prim1, z1, nc, fc, dt=dt, checkpoint_every=min(ck, nc))
    fd_short = fd_checks(nc, fc, zbc, s0c)
    out = {
        "metric": "adjoint/forward time ratio (gauge-misfit gradient wrt initial state and bathymetry)",
        "value": ms[1] / ms_fwd, "unit": "ratio", "higher_is_better": False,
        "gradient_cost_ratio": (ms[0] + ms[1]) / ms_fwd,
        "forward_ms": ms_fwd, "adjoint_forward_sweep_ms": ms[0], "adjoint_reverse_sweep_ms": ms[1],
        # the reverse sweep = recomputation of the segments from their
        # checkpoints (forward work) + the reverse stages proper
        "reverse_sweep_recompute_ms": split["recompute_ms"],
        "reverse_sweep_reverse_stages_ms": split["reverse_stages_ms"],
        "reverse_stages_per_forward": split["reverse_stages_ms"] / ms_fwd,
        "forward_cell_updates_per_s": B * T1 * N / (ms_fwd * 1e-3),
        "loss": float(loss.item()), "member0_slice_bitwise": slice_same,
        "tangent_duality": tangent, "fd_full_horizon": fd_full, "fd_short_horizon": fd_short,
        "config": {"workload": f"bathymetry_inversion_batch{B}_T{T1}", "members": B, "cells": B * T1,
                   "steps": N, "dt": dt, "gauges_per_member": int(len(g0)), "checkpoint_every": ck,
                   "scheme": "SSPRK2", "flux": "HLLC<Einfeldt>", "gradient": "torch.autograd.Function",
                   "fused_stages": fused},
        "dtype": "f64", "data": "synthetic", "n_gpus": 1,
    }
    return out


def bench_adjoint(args):
    """The full inversion workload (2,000 steps) as one JSON line."""
    out = adjoint_measure(args.adj_batch, args.adj_n, args.adj_steps, args.adj_dt, args.adj_checkpoint,
                          args.adj_fd, args.adj_check_steps, fused=not args.no_fused)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=2828, help="jittered grid size (T = 2 n^2)")
    ap.add_argument("--kernel-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-only", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of 2 GiB x 5 (counter calibration)")
    ap.add_argument("--ensemble", action="store_true",
                    help="batched ensemble vs sequential single-mesh solves (one JSON line)")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--ensemble-cells", type=int, default=20_000_000)
    ap.add_argument("--t-end", type=float, default=1.0)
    ap.add_argument("--seq-steps", type=int, default=40)
    ap.add_argument("--chunk", type=int, default=0,
                    help="ensemble: run the batch in chunks of this many steps and report each chunk's rate")
    ap.add_argument("--adjoint", action="store_true",
                    help="bathymetry-inversion adjoint workload instead of the dam break")
    ap.add_argument("--adj-n", type=int, default=354, help="adjoint mesh size (T = 2 n^2)")
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--adj-dt", type=float, default=2e-4)
    ap.add_argument("--adj-checkpoint", type=int, default=50,
                    help="adjoint checkpoint interval (0: one segment, every step's intermediates kept)")
    ap.add_argument("--adj-check-steps", type=int, default=100,
                    help="horizon of the short finite-difference gradient checks")
    ap.add_argument("--adj-batch", type=int, default=32, help="inversions in the batch")
    ap.add_argument("--adj-fd", type=int, default=16, help="finite-difference directions")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default work items per thread (profiling runs)")
    ap.add_argument("--no-fused", action="store_true",
                    help="separate flux, draining and update kernels in every stage (A/B runs)")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the dam break (no ensemble, adjoint and CPU-row sub-objects)")
    ap.add_argument("--ens-t-end", type=float, default=0.2,
                    help="t_end of the default line's shortened ensemble run")
    ap.add_argument("--adj-short-steps", type=int, default=200,
                    help="steps of the default line's shortened adjoint run")
    args = ap.parse_args()
    if args.ensemble:
        bench_ensemble(args)
        return
    if args.adjoint:
        bench_adjoint(args)
        return
    if args.copy_calibration:
        import torch
        a = torch.ones(2 ** 28, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 2 * a.numel() * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import swe_fvm_amd as swe
    from swe_fvm_amd import workloads as W

    t0 = time.time()
    dom, prim = W.dam_break(n=args.n, seed=3)
    T = prim.shape[0]
    E = dom.topology.NumEdges()
    V = dom.topology.NumNodes()
    print(f"[bench] mesh+state {T} cells {E} edges {V} nodes in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    workload = f"dam_break_T{T}"
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dom, prim, "HLLC", "Einfeldt")
        gc.collect()
        print(f"[bench] cpu baseline {cpu['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
    if args.cpu_only:
        print(json.dumps({"cpu_baseline": cpu}))
        return

    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim, device=0)
    fused = sd.set_fused(not args.no_fused)
    if args.warmup:
        sd.run("SSPRK2", args.warmup, dt=0.0, cfl=0.15)
    # work items per thread: the fastest bit-identical kernel instantiation
    tn = sd.tuning()
    tune = {"flux_items": tn[0], "update_items": tn[1], "flux_block": tn[2], "update_block": tn[3],
            "layout": sd.layout(), "stage_ms": None} \
        if args.no_autotune else sd.autotune_layout(reps=5)
    print(f"[bench] autotune: flux x{tune['flux_items']}/{tune['flux_block']} "
          f"update x{tune['update_items']}/{tune['update_block']} layout {layout_name(tune['layout'])} "
          f"(stage ms per layout {tune['stage_ms']})", file=sys.stderr, flush=True)
    sd.synchronize()
    t1 = time.perf_counter()
    ms = sd.time_run("SSPRK2", args.steps, dt=0.0, cfl=0.15)
    sd.synchronize()
    wall = time.perf_counter() - t1
    ms_per_step = ms / args.steps
    value = T * args.steps / (ms * 1e-3)
    print(f"[bench] {args.steps} steps {ms:.1f} ms (wall {wall * 1e3:.1f} ms)", file=sys.stderr, flush=True)

    # per-kernel average launch durations (events bracket each launch)
    kms_total, kms = sd.time_run("SSPRK2", args.kernel_steps, dt=0.0, cfl=0.15, per_kernel=True)
    own = kernel_own_bytes(T, E, V)
    # launches per SSP-RK2 step: two reconstructions; in stage 1 (CFL mode)
    # the separate flux kernel, then the tile kernel reading its fluxes
    # (fused_ff), in stage 2 the fused flux-draining-update tile kernel
    # (without tiles: flux, draining and update kernels in both stages)
    per_step = {"recon1": 2, "recon2": 2, "flux": 1 if fused else 2, "drain": 2, "update": 2, "fused": 1,
                "fused_ff": 1}
    names = ["recon1", "recon2", "flux", "drain", "update", "fused", "fused_ff"]
    per = {k: {"avg_ms": float(kms[q]), "launches_per_step": per_step[k], "own_array_bytes": own[k],
               "own_array_gbs": own[k] / (kms[q] * 1e-3) / 1e9}
           for q, k in enumerate(names) if kms[q] > 0}
    # the dominant kernel: most time per step; its achieved bandwidth counts
    # the compulsory bytes of the stage it completes (the fused kernel and
    # the stage-2 update: 108 B per cell; a stage-1 kernel: 84 B) against
    # its own launch time
    dom_k = max(per, key=lambda k: per[k]["avg_ms"] * per[k]["launches_per_step"])
    stage_b = STAGE_BYTES[1] if dom_k == "fused" else STAGE_BYTES[0]
    achieved = stage_b * T / (per[dom_k]["avg_ms"] * 1e-3) / 1e9
    pmc = load_traffic(workload)
    traffic = pmc.get(dom_k, {}).get("bytes_per_launch") if pmc.get("fused_stages", False) == fused else None
    step_traffic = None
    if traffic is not None:
        step_traffic = sum(pmc[k]["bytes_per_launch"] * per_step[k] for k in per if k in pmc)
    step_gbs = STEP_BYTES_PER_CELL * value / 1e9
    h = sd.GetVolField()
    finite = bool(np.isfinite(h).all())
    del sd
    gc.collect()
    print(f"[bench] dominant kernel {dom_k}: {per[dom_k]['avg_ms']:.3f} ms, {achieved:.0f} GB/s compulsory",
          file=sys.stderr, flush=True)

    fl_rows = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    extras = {}
    if not args.no_extras:
        t0 = time.time()
        extras["ensemble"] = ensemble_measure(256, 20_000_000, args.ens_t_end, 10, host_stepped=False, tune_reps=2)
        extras["ensemble"]["note"] = (f"the 256-member ensemble run shortened to t_end = {args.ens_t_end} (the full run to t = 1.0: "
                                      f"python bench.py --ensemble)")
        gc.collect()
        print(f"[bench] ensemble {extras['ensemble']['value']:.2f}x in {time.time() - t0:.0f}s",
              file=sys.stderr, flush=True)
        t0 = time.time()
        adj = adjoint_measure(B=32, n=354, N=args.adj_short_steps, dt=2e-4, checkpoint=50, nfd=4,
                              check_steps=50)
        for key in ("fd_full_horizon", "fd_short_horizon", "tangent_duality"):
            adj[key].pop("results", None)
        adj["note"] = (f"the 32-member inversion shortened to {args.adj_short_steps} steps, checkpoints every 50 "
                       f"(the full 2,000 steps: python bench.py --adjoint)")
        extras["adjoint"] = adj
        gc.collect()
        print(f"[bench] adjoint/forward {adj['value']:.2f} in {time.time() - t0:.0f}s", file=sys.stderr,
              flush=True)
        if cpu is not None:
            t0 = time.time()
            extras["cpu_rows"] = cpu_rows(fl_rows)
            print(f"[bench] cpu rows in {time.time() - t0:.0f}s", file=sys.stderr, flush=True)
    out = {
        "metric": "FP64 cell-updates/s/GPU",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "cells": T, "edges": E, "nodes": V,
                   "mesh": f"seeded Delaunay of {args.n}x{args.n} jittered grid, Morton order",
                   "scheme": "SSPRK2", "flux": "HLLC<Einfeldt>", "dt": "CFL 0.15",
                   "parallelism": "single GPU", "fused_stages": fused,
                   "tuning": {"flux_edges_per_thread": tune["flux_items"],
                              "update_cells_per_thread": tune["update_items"],
                              "flux_block_threads": tune["flux_block"],
                              "update_block_threads": tune["update_block"],
                              "state_layout": layout_name(tune["layout"]),
                              "stage_ms_per_layout": tune["stage_ms"]}},
        "roofline": {"bound": "hbm", "kernel": dom_k, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_cell": stage_b,
                     "basis": "compulsory bytes of the stage the kernel completes per cell (SSP-RK2 stage "
                              "1: 84 B, stage 2: 108 B) / the kernel's average launch time"},
        "step_roofline": {"bytes_per_cell_update": STEP_BYTES_PER_CELL, "achieved": step_gbs,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "traffic_per_step": step_traffic,
                          "traffic_ratio": step_traffic / (STEP_BYTES_PER_CELL * T) if step_traffic else None},
        "kernels": per,
        "cpu_baseline": cpu,
        "finite": finite,
    }
    out.update(extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
