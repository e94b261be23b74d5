"""Pointwise model functions for the oracle (test infrastructure only).

Each model restates, in numpy, the functions text2code generates from a
pdemodel file (the ``Flux``, ``Source`` and ``FbouHdg`` entries), with their
Jacobians written out by hand:

* ``Poisson``      -- examples/Poisson/poisson2d/pdemodel.txt and
                      examples/Poisson/poisson3d/pdemodel.txt
* ``PoissonPoly``  -- the same flux with the exact solution u = |x|^2, used to pin
                      the discretisation (HDG reproduces polynomials of degree <= p)
* ``ConvDiff``     -- steady convection-diffusion, flux kappa*q + c*u per
                      component, source read from the auxiliary field v (odg)

Array conventions (numpy, point index first):
  x[N, nd], uq[N, nc], v[N, nco], uh[N, ncu], nl[N, nd]
  flux     -> f[N, ncu, nd],  f_uq[N, ncu, nd, nc]
  source   -> s[N, ncu],      s_uq[N, ncu, nc]
  fbou(ib) -> fb[N, ncu],     fb_uq[N, ncu, nc],  fb_uh[N, ncu, ncu]
``ib`` is the 1-based boundary-condition number (uequation.cpp:392-398 passes ibc+1).
"""
import numpy as np


class Poisson:
    """Flux kappa*q, source 2 pi^2 sin(pi x) sin(pi y) (3 pi^2 ... in 3D), FbouHdg tau(0-uh)."""
    name = "poisson"

    def __init__(self, nd, ncu=1):
        self.nd, self.ncu = nd, ncu
        self.nc = ncu * (nd + 1)

    def flux(self, x, uq, v, param, time):
        N, nd, ncu, nc = x.shape[0], self.nd, self.ncu, self.nc
        kappa = param[0]
        f = np.zeros((N, ncu, nd))
        f_uq = np.zeros((N, ncu, nd, nc))
        for c in range(ncu):
            for d in range(nd):
                qi = ncu + c + ncu * d
                f[:, c, d] = kappa * uq[:, qi]
                f_uq[:, c, d, qi] = kappa
        return f, f_uq

    def exact(self, x):
        return np.prod(np.sin(np.pi * x), axis=1)

    def source(self, x, uq, v, param, time):
        N = x.shape[0]
        s = np.zeros((N, self.ncu))
        s[:, 0] = self.nd * np.pi ** 2 * np.prod(np.sin(np.pi * x), axis=1)
        return s, np.zeros((N, self.ncu, self.nc))

    def boundary_value(self, x):
        return np.zeros(x.shape[0])

    def fbou(self, ib, x, uq, v, uh, nl, tau, param, time):
        N, ncu, nc = x.shape[0], self.ncu, self.nc
        g = self.boundary_value(x)
        fb = np.zeros((N, ncu))
        fb_uh = np.zeros((N, ncu, ncu))
        for c in range(ncu):
            fb[:, c] = tau[0] * (g - uh[:, c])
            fb_uh[:, c, c] = -tau[0]
        return fb, np.zeros((N, ncu, nc)), fb_uh


class PoissonPoly(Poisson):
    """Exact solution u = sum_d x_d^2, so -kappa*Laplace(u) = -2 nd kappa; Dirichlet data u."""
    name = "poisson_poly"

    def exact(self, x):
        return np.sum(x * x, axis=1)

    def source(self, x, uq, v, param, time):
        N = x.shape[0]
        s = np.zeros((N, self.ncu))
        s[:, 0] = -2.0 * self.nd * param[0]
        return s, np.zeros((N, self.ncu, self.nc))

    def boundary_value(self, x):
        return self.exact(x)


class ConvDiff:
    """f[c, d] = kappa*q[c, d] + vel[d]*u[c]; s[c] = v[c];
    FbouHdg: ib == 1 Dirichlet tau*(0 - uh), otherwise homogeneous Neumann
    (zero diffusive flux) kappa*q.n + tau*(u - uh).  physicsparam = [kappa, vel_1..vel_nd]."""
    name = "convdiff"

    def __init__(self, nd, ncu):
        self.nd, self.ncu = nd, ncu
        self.nc = ncu * (nd + 1)

    def flux(self, x, uq, v, param, time):
        N, nd, ncu, nc = x.shape[0], self.nd, self.ncu, self.nc
        kappa = param[0]
        vel = param[1:1 + nd]
        f = np.zeros((N, ncu, nd))
        f_uq = np.zeros((N, ncu, nd, nc))
        for c in range(ncu):
            for d in range(nd):
                qi = ncu + c + ncu * d
                f[:, c, d] = kappa * uq[:, qi] + vel[d] * uq[:, c]
                f_uq[:, c, d, qi] = kappa
                f_uq[:, c, d, c] = vel[d]
        return f, f_uq

    def source(self, x, uq, v, param, time):
        N = x.shape[0]
        return v[:, :self.ncu].copy(), np.zeros((N, self.ncu, self.nc))

    def fbou(self, ib, x, uq, v, uh, nl, tau, param, time):
        N, nd, ncu, nc = x.shape[0], self.nd, self.ncu, self.nc
        fb = np.zeros((N, ncu))
        fb_uq = np.zeros((N, ncu, nc))
        fb_uh = np.zeros((N, ncu, ncu))
        if ib == 1:
            for c in range(ncu):
                fb[:, c] = tau[0] * (0.0 - uh[:, c])
                fb_uh[:, c, c] = -tau[0]
            return fb, fb_uq, fb_uh
        kappa = param[0]
        for c in range(ncu):
            acc = tau[0] * (uq[:, c] - uh[:, c])
            for d in range(nd):
                qi = ncu + c + ncu * d
                acc = acc + kappa * uq[:, qi] * nl[:, d]
                fb_uq[:, c, qi] = kappa * nl[:, d]
            fb[:, c] = acc
            fb_uq[:, c, c] = tau[0]
            fb_uh[:, c, c] -= tau[0]
        return fb, fb_uq, fb_uh


def make_model(name, nd, ncu):
    if name == "poisson":
        return Poisson(nd, ncu)
    if name == "poisson_poly":
        return PoissonPoly(nd, ncu)
    if name == "convdiff":
        return ConvDiff(nd, ncu)
    # a model defined by a pdemodel file: exasim_amd/models/<name>.txt, interpreted
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "exasim_amd", "models", name + ".txt")
    if os.path.exists(path):
        from .dslmodel import DSLModel
        return DSLModel(path, nd, ncu)
    raise ValueError("unknown model " + name)
