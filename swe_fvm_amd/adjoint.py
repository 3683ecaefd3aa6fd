"""Gradients of functionals of the final state with respect to the initial
state and the node (or per-cell) bathymetry, through the hand-written adjoint of the step
(swe_adjoint_run): a torch.autograd.Function and the gauge-misfit loss of
the bathymetry-inversion workload.

The forward of the Function runs the steps on the device; its backward
re-runs them with checkpoints and sweeps the adjoint back (step sizes held
fixed: a fixed dt, or the sizes the forward's CFL run took).  Tensors are
float64 on the CPU; the state lives on the GPU in between.
"""
import numpy as np

try:
    import torch
except ImportError:  # the numpy interface below works without torch
    torch = None


class CellBathymetry:
    """Per-cell bathymetry b (T values) on a mesh whose solver reads node
    heights z (V values; the reference's Domain holds node z,
    include/Bathymetry.h:43, and a cell's bathymetry is its centroid mean,
    src/ValueField.cpp:8-10): z = M b with z_n the mean of b over the cells
    incident to node n.  The gradient of a function of z with respect to b
    is M^T dJ/dz."""

    def __init__(self, tp, num_nodes):
        tp = np.asarray(tp, dtype=np.int64).reshape(-1, 3)
        self.T, self.V = tp.shape[0], int(num_nodes)
        self.rows = tp.reshape(-1)                      # node of each corner
        self.cols = np.repeat(np.arange(self.T), 3)     # cell of each corner
        deg = np.bincount(self.rows, minlength=self.V).astype(np.float64)
        if np.any(deg == 0):
            raise ValueError("a node belongs to no cell")
        self.inv_deg = 1.0 / deg

    def to_nodes(self, b):
        """z = M b"""
        b = np.asarray(b, dtype=np.float64).reshape(self.T)
        return np.bincount(self.rows, weights=b[self.cols], minlength=self.V) * self.inv_deg

    def gradient(self, zbar):
        """M^T zbar: the per-cell gradient from the node one"""
        zbar = np.asarray(zbar, dtype=np.float64).reshape(self.V)
        return np.bincount(self.cols, weights=(zbar * self.inv_deg)[self.rows], minlength=self.T)


def run_and_adjoint(sd, prim0, z, nsteps, final_bar_fn, scheme="SSPRK2", dt=None, dts=None,
                    checkpoint_every=50, cell_bathymetry=None):
    """Sets (prim0, z), runs nsteps steps, evaluates final_bar_fn(final prim)
    -> (J, dJ/dprim0, dJ/dz, [forward ms, reverse ms]).  With
    cell_bathymetry (a CellBathymetry of the mesh) z holds per-cell values
    (node heights M z) and the gradient returned is the per-cell one."""
    if cell_bathymetry is not None:
        J, s0, zb, ms = run_and_adjoint(sd, prim0, cell_bathymetry.to_nodes(z), nsteps, final_bar_fn,
                                        scheme, dt, dts, checkpoint_every)
        return J, s0, cell_bathymetry.gradient(zb), ms
    sd.set_bathymetry(z)
    sd.SetVolField(prim0)
    # forward pass for the loss value (the adjoint sweep repeats it)
    if dts is None:
        sd.run(scheme, nsteps, dt=dt)
    else:
        for d in dts:
            sd.run(scheme, 1, dt=float(d))
    J, gbar = final_bar_fn(sd.GetVolField())
    sd.SetVolField(prim0)
    s0, zb, ms = sd.adjoint_run(gbar, nsteps, scheme=scheme, dt=dt, dts=dts, checkpoint_every=checkpoint_every)
    return J, s0, zb, ms


def batch_gauge_misfit(members):
    """Sum over batch members of their gauge misfits: members is a list of
    (gauge cell ids in the stacked domain, targets); returns fn like
    gauge_misfit and the per-member values of the last call in fn.parts."""
    fns = [gauge_misfit(g, t) for g, t in members]

    def fn(prim):
        g = np.zeros_like(prim)
        parts = []
        for f in fns:
            j, gk = f(prim)
            parts.append(j)
            g += gk
        fn.parts = parts
        return float(np.sum(parts)), g
    fn.parts = []
    return fn


def gauge_misfit(gauges, target):
    """J = sum over gauge cells of (w - target)^2 and its adjoint."""
    gauges = np.asarray(gauges)
    target = np.asarray(target, dtype=np.float64)

    def fn(prim):
        r = prim[gauges, 0] - target
        g = np.zeros_like(prim)
        np.add.at(g[:, 0], gauges, 2.0 * r)
        return float(np.sum(r * r)), g
    return fn


if torch is not None:
    class ShallowWaterRun(torch.autograd.Function):
        """final_prim = ShallowWaterRun.apply(prim0, z, sd, nsteps, scheme, dt,
        checkpoint_every, cells): prim0 (T x 3) and z float64 CPU tensors; the
        backward gives their gradients through the adjoint of every stage.
        z holds the node heights (V), or with cells (a CellBathymetry of the
        mesh) per-cell bathymetry values (T) mapped to the nodes.  dt <= 0
        runs in CFL mode and the backward uses the step sizes it took.  sd
        is a swe_fvm_amd.SpaceDisc (ctypes) or a Solver of the pybind module
        SWE_FVM; both call the same C ABI."""

        @staticmethod
        def forward(ctx, prim0, z, sd, nsteps, scheme="SSPRK2", dt=1e-3, checkpoint_every=50, cells=None):
            p0 = prim0.detach().cpu().numpy().astype(np.float64)
            zz = z.detach().cpu().numpy().astype(np.float64)
            ctx.cells = cells
            if cells is not None:
                zz = cells.to_nodes(zz)
            sd.set_bathymetry(zz)
            sd.SetVolField(p0)
            sd.run(scheme, nsteps, dt=dt)
            out = torch.from_numpy(sd.GetVolField().copy())
            ctx.sd, ctx.p0, ctx.z = sd, p0, zz
            # a CFL run (dt <= 0) is differentiated with the step sizes it took
            ctx.dts = None if (dt is not None and dt > 0) else sd.dts(nsteps)
            ctx.args = (nsteps, scheme, dt, checkpoint_every)
            return out

        @staticmethod
        def backward(ctx, gout):
            nsteps, scheme, dt, ck = ctx.args
            sd = ctx.sd
            sd.set_bathymetry(ctx.z)
            sd.SetVolField(ctx.p0)
            s0, zb, ms = sd.adjoint_run(gout.detach().cpu().numpy().astype(np.float64), nsteps, scheme=scheme,
                                        dt=None if ctx.dts is not None else dt, dts=ctx.dts,
                                        checkpoint_every=ck)
            sd.last_adjoint_ms = ms  # forward-sweep and reverse-sweep milliseconds
            if ctx.cells is not None:
                zb = ctx.cells.gradient(zb)
            return torch.from_numpy(s0), torch.from_numpy(zb), None, None, None, None, None, None
