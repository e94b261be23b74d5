"""Pins of the oracle's BDF time stepping and of the Tdfunc term (CPU only).

* BDF coefficients -> the values the reference tabulates in BDFcoeff (timestepcoeff.cpp:106-130)
  and the order conditions sum c_k = 0, sum c_k (-k) = 1 (consistency of u_t)
* history, stage source and time term -> the observed order of accuracy against the exact
  decay of the heat equation, exp(-2 pi^2 t) sin(pi x) sin(pi y): 1 for BDF1, 2 for BDF2 and at
  least 2 for BDF3 (its first step is a BDF1 step, whose O(dt^2) error stays in the result),
  and BDF3's error a tenth of BDF1's or less
* Tdfunc -> (c0 u_t = kappa lap u, stabilization tau) gives the solution of
  (u_t = (kappa / c0) lap u, stabilization tau / c0), the same equations divided by c0: the
  capacity must multiply the time term in the residual and in the Jacobian
"""
import numpy as np
import pytest

from exasim_amd import cases
from oracle import hdg as ora

REFERENCE_TABLE = {1: [1.0, -1.0], 2: [1.5, -2.0, 0.5], 3: [11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0]}


@pytest.mark.parametrize("order", [1, 2, 3])
def test_bdf_coefficients(order):
    c = np.array(ora.bdf_coeffs(order))
    assert np.abs(c - REFERENCE_TABLE[order]).max() < 1e-15
    k = np.arange(order + 1)
    assert abs(c.sum()) < 1e-14                       # constants have no time derivative
    assert abs((c * (-k)).sum() - 1.0) < 1e-14        # u_t of a linear function


@pytest.mark.parametrize("order", [1, 2, 3])
def test_bdf_order_of_accuracy(tmp_path, order):
    T = 0.1
    errs = []
    for k, nsteps in enumerate([10, 20]):
        path = str(tmp_path / ("s%d" % k))
        cases.convdiff_unsteady(path, nd=2, n=4, porder=5, dt=T / nsteps, nsteps=nsteps, torder=order,
                                scheme=1)
        P = ora.HDGProblem(path, "convdiff")
        assert P.tscheme == 1
        ora.time_solve(P)
        exact = np.exp(-2 * np.pi ** 2 * T) * np.prod(np.sin(np.pi * P.xdg), axis=2)
        errs.append(np.abs(P.udg[:, :, 0] - exact).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > min(order, 2) - 0.3, (errs, rate)
    ERRS[order] = errs[1]
    # the self-started BDF3 (one BDF1 and one BDF2 step first) keeps the start-up's second-order
    # error, so it is checked against the first-order scheme only
    if order == 3 and 1 in ERRS:
        assert ERRS[3] < 0.1 * ERRS[1]


ERRS = {}


@pytest.mark.parametrize("scheme", [0, 1])
def test_tdfunc_capacity_scales_time(tmp_path, scheme):
    kappa, c0 = 0.8, 2.0
    a, b = str(tmp_path / "cap"), str(tmp_path / "ref")
    kw = dict(nd=2, n=3, porder=3, dt=0.02, nsteps=3, torder=2, nstage=2, scheme=scheme, newton_iter=3)
    cases.convdiff_unsteady(a, kappa=kappa, capacity=(c0, 0.0), **kw)
    cases.convdiff_unsteady(b, kappa=kappa / c0, tau=1.0 / c0, **kw)
    P = ora.HDGProblem(a, "heat2d_tdfunc")
    assert P.tdfunc == 1
    Q = ora.HDGProblem(b, "convdiff")
    ora.time_solve(P)
    ora.time_solve(Q)
    assert np.abs(P.udg[:, :, 0] - Q.udg[:, :, 0]).max() < 1e-11 * np.abs(Q.udg[:, :, 0]).max()
    # with the flag off the capacity is ignored
    c = str(tmp_path / "off")
    cases.convdiff_unsteady(c, kappa=kappa, capacity=(c0, 0.0), **kw)
    R = ora.HDGProblem(c, "heat2d_tdfunc")
    R.tdfunc = 0
    ora.time_solve(R)
    assert np.abs(R.udg[:, :, 0] - P.udg[:, :, 0]).max() > 1e-4
