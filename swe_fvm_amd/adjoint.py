"""Gradients of functionals of the final state with respect to the initial
state and the node bathymetry, through the hand-written adjoint of the step
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


def run_and_adjoint(sd, prim0, z, nsteps, final_bar_fn, scheme="SSPRK2", dt=None, dts=None,
                    checkpoint_every=50):
    """Sets (prim0, z), runs nsteps steps, evaluates final_bar_fn(final prim)
    -> (J, dJ/dfinal), and returns (J, dJ/dprim0, dJ/dz, [forward ms,
    reverse ms])."""
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
        checkpoint_every): prim0 (T x 3) and z (V) float64 CPU tensors; the
        backward gives their gradients through the adjoint of every stage.
        sd is a swe_fvm_amd.SpaceDisc (ctypes) or a Solver of the pybind
        module SWE_FVM; both call the same C ABI."""

        @staticmethod
        def forward(ctx, prim0, z, sd, nsteps, scheme="SSPRK2", dt=1e-3, checkpoint_every=50):
            p0 = prim0.detach().cpu().numpy().astype(np.float64)
            zz = z.detach().cpu().numpy().astype(np.float64)
            sd.set_bathymetry(zz)
            sd.SetVolField(p0)
            sd.run(scheme, nsteps, dt=dt)
            out = torch.from_numpy(sd.GetVolField().copy())
            ctx.sd, ctx.p0, ctx.z = sd, p0, zz
            ctx.args = (nsteps, scheme, dt, checkpoint_every)
            return out

        @staticmethod
        def backward(ctx, gout):
            nsteps, scheme, dt, ck = ctx.args
            sd = ctx.sd
            sd.set_bathymetry(ctx.z)
            sd.SetVolField(ctx.p0)
            s0, zb, ms = sd.adjoint_run(gout.detach().cpu().numpy().astype(np.float64), nsteps, scheme=scheme,
                                        dt=dt, checkpoint_every=ck)
            sd.last_adjoint_ms = ms  # forward-sweep and reverse-sweep milliseconds
            return torch.from_numpy(s0), torch.from_numpy(zb), None, None, None, None, None
