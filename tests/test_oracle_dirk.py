"""Pins of the oracle's DIRK time stepping (CPU only).

* coefficients -> the d, c, t values the reference tabulates in DIRKcoeff
  (timestepcoeff.cpp:35-104), here derived from the Butcher tableaux
* stage source, time term and stage combination -> the observed order of accuracy
  against the exact decay of the heat equation, exp(-2 pi^2 t) sin(pi x) sin(pi y),
  for backward Euler (1), DIRK(2,2) (2) and DIRK(3,3) (3)
"""
import numpy as np
import pytest

from exasim_amd import cases
from oracle import hdg as ora

REFERENCE_TABLE = {
    (1, 1): ([1.0], [1.0], [1.0]),
    (1, 2): ([2.0], [2.0], [0.5]),
    (2, 2): ([3.414213562373096, -8.242640687119289, 0.0, 3.414213562373096], [0.0, 1.0],
             [0.292893218813452, 1.0]),
    (2, 3): ([1.267949192431123, 0.928203230275509, 0.0, 1.267949192431123],
             [1.098076211353316, 0.633974596215561], [0.788675134594813, 0.211324865405187]),
    (3, 3): ([2.294280360323567, -1.484721005721436, -8.556127801733982, 0.0, 2.294280360323568,
              3.391748836909791, 0.0, 0.0, 2.294280360323567], [0.0, 0.0, 1.0],
             [0.4358665215, 0.71793326075, 1.0]),
    (3, 4): ([0.935822227524088, 0.497940606760015, -0.345865915800969, 0.0, 0.935822227524088,
              2.867525668568206, 0.0, 0.0, 0.935822227524088],
             [0.445622407287714, 1.064177772475912, 0.120614758428183],
             [1.068579021301629, 0.5, -0.068579021301629]),
}


@pytest.mark.parametrize("key", list(REFERENCE_TABLE))
def test_dirk_coefficients_match_reference_table(key):
    d_ref, c_ref, t_ref = REFERENCE_TABLE[key]
    c, d, t = ora.dirk_coeffs(*key)
    assert np.abs(d - d_ref).max() < 1e-13
    assert np.abs(c - c_ref).max() < 1e-13
    assert np.abs(t - t_ref).max() < 1e-13


@pytest.mark.parametrize("nstage,torder", [(1, 1), (2, 2), (3, 3)])
def test_dirk_order_of_accuracy(tmp_path, nstage, torder):
    T = 0.1
    errs = []
    for k, nsteps in enumerate([5, 10]):
        path = str(tmp_path / ("s%d" % k))
        cases.convdiff_unsteady(path, nd=2, n=4, porder=5, dt=T / nsteps, nsteps=nsteps, torder=torder,
                                nstage=nstage)
        P = ora.HDGProblem(path, "convdiff")
        ora.dirk_solve(P)
        exact = np.exp(-2 * np.pi ** 2 * T) * np.prod(np.sin(np.pi * P.xdg), axis=2)
        errs.append(np.abs(P.udg[:, :, 0] - exact).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > torder - 0.2, (errs, rate)
