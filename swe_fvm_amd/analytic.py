"""Analytic solutions and error norms (examples/Tests.h:9-280; the commented
Compare / ComputeIntegrals of src/SpaceDisc.cpp:77-138).

Every solution is in the solver's non-dimensional units (g = 1) and takes
its parameters explicitly instead of from a parsed .ini file; config.py maps
a reference-style configuration onto them.  Functions are vectorised over
numpy arrays of x and y.

Sources: the shallow-water equations these solve are
    h_t + (hu)_x + (hv)_y = 0
    (hu)_t + (hu^2 + h^2/2)_x + (huv)_y = -h b_x + cor h v - tau h u
    (hv)_t + (huv)_x + (hv^2 + h^2/2)_y = -h b_y - cor h u - tau h v
(the solver's Coriolis term, src/SpaceDisc.cpp:26-28 + src/TimeDisc.cpp:35-37;
its friction tau is stored but unused, include/SpaceDisc.h:47).
"""
import math

import numpy as np

TOL = 1e-12  # include/Bathymetry.h:5


class Analytic:
    """Test (examples/Tests.h:9-31): b, u, v, h and the derived w, hu, hv."""

    def __init__(self, mid=(0.0, 0.0), cor=0.0, tau=0.0):
        self.mx, self.my = mid
        self.cor, self.tau = cor, tau

    def b(self, x, y):
        raise NotImplementedError

    def u(self, x, y, t):
        raise NotImplementedError

    def v(self, x, y, t):
        raise NotImplementedError

    def h(self, x, y, t):
        raise NotImplementedError

    def w(self, x, y, t):
        return self.h(x, y, t) + self.b(x, y)

    def hu(self, x, y, t):
        return self.h(x, y, t) * self.u(x, y, t)

    def hv(self, x, y, t):
        return self.h(x, y, t) * self.v(x, y, t)

    def is_wet(self, x, y, t):
        return self.h(x, y, t) >= TOL

    def bath(self, x, y):  # name used by the workload generators
        return self.b(x, y)


class LakeAtRest(Analytic):
    """LakeAtRestTest (examples/Tests.h:33-45): a raised plateau on (1,3)^2."""

    def b(self, x, y):
        inside = (1.0 < x) & (x < 3.0) & (1.0 < y) & (y < 3.0)
        return np.where(inside, -0.2, -1.0)

    def u(self, x, y, t):
        return np.zeros_like(np.asarray(x, dtype=float))

    v = u

    def h(self, x, y, t):
        return np.maximum(0.0, -self.b(x, y))


class Bowl(Analytic):
    """BowlTest (examples/Tests.h:48-58): b = delta (|x - mid|^2 - 1)."""

    def __init__(self, delta=1.0, **kw):
        super().__init__(**kw)
        self.delta = delta

    def b(self, x, y):
        return self.delta * ((x - self.mx) * (x - self.mx) + (y - self.my) * (y - self.my) - 1.0)


class Ball(Bowl):
    """BallTest (examples/Tests.h:60-133): a lens of water oscillating in the
    bowl, with damping tau and rotation cor; (x0, y0) offset of its centre
    from mid, (u0, v0) its initial velocity."""

    def __init__(self, x0=0.2, y0=0.0, u0=0.0, v0=0.0, **kw):
        super().__init__(**kw)
        d, c, ta = self.delta, self.cor, self.tau
        self.x0, self.y0, self.u0, self.v0 = x0, y0, u0, v0
        lam1 = 2.0 * d - 0.25 * ta * ta + 0.25 * c * c
        lam2 = 0.5 * c * ta
        root = lam1 + math.sqrt(lam1 * lam1 + lam2 * lam2)
        self.beta = 0.5 * math.sqrt(2.0 * root)
        self.alpha = lam2 / math.sqrt(2.0 * root)
        al, be = self.alpha, self.beta
        s = al * al + be * be
        k = al * al + be * be + 0.5 * al * ta + 0.5 * be * c
        m = be * ta - al * c
        self.a1 = (0.5 * k * x0 + 0.25 * m * y0 + 0.5 * (al * u0 + be * v0)) / s
        self.a2 = (0.5 * k * y0 - 0.25 * m * x0 + 0.5 * (al * v0 - be * u0)) / s
        self.a3 = x0 - self.a1
        self.a4 = y0 - self.a2

    def _modes(self, t):
        c, ta, al, be = self.cor, self.tau, self.alpha, self.beta
        em, ep = math.exp(-(0.5 * ta - al) * t), math.exp(-(0.5 * ta + al) * t)
        cm, sm = math.cos((0.5 * c - be) * t), math.sin((0.5 * c - be) * t)
        cp, sp = math.cos((0.5 * c + be) * t), math.sin((0.5 * c + be) * t)
        phim = em * (self.a1 * cm + self.a2 * sm)
        phip = ep * (self.a3 * cp + self.a4 * sp)
        psim = em * (self.a2 * cm - self.a1 * sm)
        psip = ep * (self.a4 * cp - self.a3 * sp)
        return phim, phip, psim, psip

    def _vel(self, t):
        c, ta, al, be = self.cor, self.tau, self.alpha, self.beta
        phim, phip, psim, psip = self._modes(t)
        u = (0.5 * c + be) * psip - (0.5 * ta + al) * phip + (0.5 * c - be) * psim - (0.5 * ta - al) * phim
        v = -(0.5 * c + be) * phip - (0.5 * ta + al) * psip - (0.5 * c - be) * phim - (0.5 * ta - al) * psim
        return u, v

    def u(self, x, y, t):
        return np.where(self.is_wet(x, y, t), self._vel(t)[0], 0.0)

    def v(self, x, y, t):
        return np.where(self.is_wet(x, y, t), self._vel(t)[1], 0.0)

    def h(self, x, y, t):
        phim, phip, psim, psip = self._modes(t)
        fx = x - self.mx - phip - phim
        fy = y - self.my - psip - psim
        return np.maximum(0.0, self.delta * (1.0 - fx * fx - fy * fy))


class Thacker(Bowl):
    """ThackerTest (examples/Tests.h:136-163): u = p(t) dx + q(t) dy, a
    quadratic depth Hc + Hxx dx^2 / 2 + Hxy dx dy + Hyy dy^2 / 2."""

    def __init__(self, H0=1.25, p0=0.0, q0=0.0, **kw):
        super().__init__(**kw)
        self.H0, self.p0, self.q0 = H0, p0, q0

    def u(self, x, y, t):
        return self.p(t) * (x - self.mx) + self.q(t) * (y - self.my)

    def h(self, x, y, t):
        dx, dy = x - self.mx, y - self.my
        r = self.Hc(t) + 0.5 * self.Hxx(t) * dx * dx + self.Hxy(t) * dx * dy + 0.5 * self.Hyy(t) * dy * dy
        return np.maximum(0.0, r)


class ConservativeThacker(Thacker):
    """ConservativeThackerTest (examples/Tests.h:165-199)."""

    def __init__(self, **kw):
        super().__init__(**kw)
        self.a = self.p0 * self.p0 + self.q0 * self.q0
        self.om = math.sqrt(0.25 * self.cor * self.cor + 2.0 * self.a + 4.0 * self.delta) + 0.5 * self.cor

    def p(self, t):
        return self.p0 * math.cos(self.om * t) + self.q0 * math.sin(self.om * t)

    def q(self, t):
        return self.q0 * math.cos(self.om * t) - self.p0 * math.sin(self.om * t)

    def v(self, x, y, t):
        return self.q(t) * (x - self.mx) + self.p(t) * (self.my - y)

    def Hc(self, t):
        return self.H0

    def Hxx(self, t):
        return -self.a - self.q(t) * (self.om - self.cor) - 2.0 * self.delta

    def Hxy(self, t):
        return self.p(t) * (self.om - self.cor)

    def Hyy(self, t):
        return -self.a + self.q(t) * (self.om - self.cor) - 2.0 * self.delta


class SolenoidalThacker(Thacker):
    """SolenoidalThackerTest (examples/Tests.h:201-235), a divergence-free
    rotating lens.  Without friction it solves the equations for p0 = 0 (any
    rotation) and for p0^2 = q0^2 - 2 delta without rotation (then Hxx = 0);
    the reference's assertion p0 = sqrt(q0 - 2 delta) (examples/Tests.h:212)
    gives no solution unless q0 = 1 or p0 = 0."""

    def p(self, t):
        return self.p0 * math.exp(-self.tau * t)

    def q(self, t):
        return self.q0 * math.exp(-self.tau * t)

    def v(self, x, y, t):
        return self.q(t) * (self.mx - x) + self.p(t) * (self.my - y)

    def Hc(self, t):
        return self.H0

    def Hxx(self, t):
        if self.tau != 0:
            return 0.0
        return self.q0 * self.q0 - self.p0 * self.p0 - self.q0 * self.cor - 2.0 * self.delta

    def Hxy(self, t):
        return self.p0 * self.cor

    Hyy = Hxx


class ClassicThacker(Thacker):
    """ClassicThackerTest (examples/Tests.h:237-280): the periodic paraboloid
    with period 2 pi / sqrt(cor^2 + 8 delta)."""

    def __init__(self, **kw):
        super().__init__(**kw)
        c, p0, q0, H0 = self.cor, self.p0, self.q0, self.H0
        self.om = math.sqrt(c * c + 8.0 * self.delta)
        qz = (q0 - 0.5 * c) ** 2
        rz = qz + 2.0 * H0 * H0 + p0 * p0 - 0.25 * self.om * self.om
        self.a = math.sqrt(rz * rz + self.om * self.om * p0 * p0) / (rz + 0.5 * self.om * self.om)
        self.phase = math.atan(self.om * p0 / rz)
        self.period = 2.0 * math.pi / self.om

    def _den(self, t):
        return 1.0 - self.a * math.cos(self.om * t + self.phase)

    def p(self, t):
        return 0.5 * self.om * self.a * math.sin(self.om * t + self.phase) / self._den(t)

    def q(self, t):
        return (self.q0 - 0.5 * self.cor) * (1.0 - self.a * math.cos(self.phase)) / self._den(t) + 0.5 * self.cor

    def v(self, x, y, t):
        return self.q(t) * (self.mx - x) + self.p(t) * (y - self.my)

    def Hc(self, t):
        return self.H0 * (1.0 - self.a * math.cos(self.phase)) / self._den(t)

    def Hxx(self, t):
        qz = (self.q0 - 0.5 * self.cor) ** 2
        az0 = (1.0 - self.a * math.cos(self.phase)) ** 2
        return (0.25 * self.om * self.om * (self.a * self.a - 1.0) + qz * az0) / self._den(t) ** 2

    def Hxy(self, t):
        return 0.0

    Hyy = Hxx


# ------------------------------------------------------------------ norms --

def triang_average(p0, p1, p2, f, n):
    """TriangAverage<k, n> (include/PointOperations.h:24-44) vectorised over
    cells: the mean of f over the n^2 sub-triangle centroids.  p0, p1, p2 are
    (T, d) vertex arrays (d = 2 or 3, points carried in d dimensions); f maps
    a (T, d) point array to a (k, T) array."""
    hs = 1.0 / n
    di = hs * (p1 - p0)
    dj = hs * (p2 - p0)
    dt = (1.0 / 3.0) * (di + dj)
    total = None
    pi = p0.copy()
    for i in range(n):
        pt = pi + dt
        for _ in range(n - i - 1):
            a = hs * np.asarray(f(pt))
            b = hs * np.asarray(f(pt + dt))
            total = (a if total is None else total + a) + b
            pt = pt + dj
        a = hs * np.asarray(f(pt))
        total = a if total is None else total + a
        pi = pi + di
    return hs * total


def compare(sd, test, t, n_partial=100):
    """Compare (src/SpaceDisc.cpp:106-138, commented out in the reference):
    per-cell mean of |(w, hu, hv)_exact - (w, hu, hv)_reconstructed| over the
    cell, by TriangAverage<3, 1> (the centroid) for dry and fully wet cells and
    <3, n_partial> for partially wet ones.  The reconstruction is the solver's
    final MUSCL of its current state (computed here on the device).  Returns
    a (T, 3) array."""
    sd.ComputeInterfaceValues()
    cls = sd.CellClass()
    w0, gx, gy = sd.GetMuscl()
    prim = sd.GetVolField()
    xyz = sd.GetDomain().geometry
    tp = sd.GetDomain().topology.tables()["tp"]
    P = [xyz[tp[:, k]] for k in range(3)]
    cen = sd.GetDomain().centroids()
    uc = np.where(cls == 0, 0.0, prim[:, 1])
    vc = np.where(cls == 0, 0.0, prim[:, 2])

    def err(sel):
        def f(q):
            x, y, z = q[:, 0], q[:, 1], q[:, 2]
            a = w0[sel] + (gx[sel] * (x - cen[sel, 0]) + gy[sel] * (y - cen[sel, 1]))
            wet = (a - z) >= 0
            a = np.where(wet, a, z)
            hq = a - z
            V = np.stack([a, hq * np.where(wet, uc[sel], 0.0), hq * np.where(wet, vc[sel], 0.0)])
            U = np.stack([test.w(x, y, t), test.hu(x, y, t), test.hv(x, y, t)])
            return np.abs(U - V)
        return f

    out = np.zeros((len(cls), 3))
    part = cls == 1
    for sel, n in ((~part, 1), (part, n_partial)):
        if sel.any():
            out[sel] = triang_average(P[0][sel], P[1][sel], P[2][sel], err(sel), n).T
    return out


def error_norms(sd, test, t, n_partial=100):
    """Area-weighted L1 and L2 norms and the maximum of the per-cell errors of
    compare(), for w, hu and hv."""
    e = compare(sd, test, t, n_partial)
    A = sd.GetDomain().areas()
    tot = A.sum()
    return {"L1": (A[:, None] * e).sum(0) / tot,
            "L2": np.sqrt((A[:, None] * e * e).sum(0) / tot),
            "Linf": e.max(0)}
