"""Pins of the oracle's w equation (wequation.cpp:31-240; CPU only).

* Chain rule: the model wsub2d (Sourcew = u - w, alpha = beta = 0, so w = u in the volume and
  w = uhat at the face points) gives the assembled system and the Newton solution of wsub2d_ref,
  the same equations with w written out, to round-off: the volume, interior-face and boundary
  terms' derivatives through w_uq are all exercised.
* Point solve: the Newton iterate satisfies the w equation, and w_uq agrees with finite
  differences of the solve.
* Time stepping: after each backward-Euler (DIRK(1,1)) or BDF1/BDF2 step the nodal w satisfies
  the discrete w equation alpha (w^{n+1} - w^n)/dt + beta w^{n+1} = Sourcew(u^{n+1}, w^{n+1})
  (BDF2: alpha (1.5 w^{n+1} - 2 w^n + 0.5 w^{n-1})/dt) with the step's nodal u.
"""
import numpy as np
import pytest

from exasim_amd import cases
from oracle import hdg as ora


def test_w_chain_rule_matches_substituted_model(tmp_path):
    a, b = str(tmp_path / "w"), str(tmp_path / "ref")
    cases.wmodel2d(a, model="wsub2d", mu=(0.5, 1.0), u0=0.3, w0=0.0)
    cases.wmodel2d(b, model="wsub2d_ref", mu=(0.5, 1.0), u0=0.3)
    P, Q = ora.HDGProblem(a, "wsub2d"), ora.HDGProblem(b, "wsub2d_ref")
    assert P.ncw == 1 and Q.ncw == 0
    for X in (P, Q):
        X.get_q(X.udg, X.uh)
    _, S1, b1 = P.assemble()
    _, S2, b2 = Q.assemble()
    for k in ("D", "F", "K", "H", "DinvF", "Ru", "Rh"):
        assert np.abs(S1[k] - S2[k]).max() <= 1e-13 * max(np.abs(S2[k]).max(), 1.0), k
    assert np.abs(b1 - b2).max() < 1e-13
    ora.newton_solve(P)
    ora.newton_solve(Q)
    assert np.abs(P.uh - Q.uh).max() < 1e-12
    assert np.abs(P.wdg[:, :, 0] - P.udg[:, :, 0]).max() < 1e-12     # GetW at the nodes: w = u


def test_w_point_solve_and_derivative(tmp_path):
    a = str(tmp_path / "w")
    cases.wmodel2d(a, model="wnl2d", mu=(1.0, 1.0, 2.0), alpha=0.0, beta=1.0, u0=0.5, w0=0.0)
    P = ora.HDGProblem(a, "wnl2d")
    P.get_q(P.udg, P.uh)
    rng = np.random.default_rng(3)
    N = 20
    x = rng.uniform(0, 1, (N, 2))
    uq = rng.uniform(-1, 1, (N, 3))
    v = np.zeros((N, 0))
    w0 = np.zeros((N, 1))
    ws = np.zeros((N, 1))
    w, w_uq = P.w_solve(x, uq, v, w0, ws)
    assert np.abs(w[:, 0] + w[:, 0] ** 3 - 2.0 * uq[:, 0]).max() < 1e-12   # beta w = mu2 u - w^3
    h = 1e-6
    for c in range(3):
        up, um = uq.copy(), uq.copy()
        up[:, c] += h
        um[:, c] -= h
        fd = (P.w_solve(x, up, v, w0, ws)[0] - P.w_solve(x, um, v, w0, ws)[0]) / (2 * h)
        assert np.abs(fd[:, 0] - w_uq[:, 0, c]).max() < 1e-7, c


@pytest.mark.parametrize("scheme,torder,nstage", [(0, 1, 1), (1, 1, 1), (1, 2, 1)])
def test_w_time_steps_satisfy_discrete_equation(tmp_path, scheme, torder, nstage):
    a = str(tmp_path / "w")
    dt, alpha, beta, mu2 = 0.05, 1.0, 1.0, 2.0
    cases.wmodel2d(a, model="wnl2d", mu=(1.0, 1.0, mu2), alpha=alpha, beta=beta, u0=0.0, w0=0.2,
                   dt=[dt] * 3, torder=torder, nstage=nstage, scheme=scheme)
    P = ora.HDGProblem(a, "wnl2d")
    hist = [P.wdg[:, :, 0].copy()]
    for step in range(3):
        P.dt = np.array([dt])
        ora.time_solve(P, nsteps=1) if scheme == 0 else None
        if scheme == 1:
            break
        u = P.udg[:, :, 0]
        w1 = P.wdg[:, :, 0]
        lhs = alpha * (w1 - hist[-1]) / dt + beta * w1
        assert np.abs(lhs - (mu2 * u - w1 ** 3)).max() < 1e-9, step
        hist.append(w1.copy())
    if scheme == 1:  # the BDF loop keeps its own history: run all steps, recording after each
        P = ora.HDGProblem(a, "wnl2d")
        ws, us = [P.wdg[:, :, 0].copy()], []
        orig = ora.newton_solve

        def record(prob, log=None):
            h = orig(prob, log)
            ws.append(prob.wdg[:, :, 0].copy())
            us.append(prob.udg[:, :, 0].copy())
            return h
        ora.newton_solve = record
        try:
            ora.bdf_solve(P)
        finally:
            ora.newton_solve = orig
        for n in range(1, len(ws)):
            if torder == 1 or n == 1:
                dwdt = (ws[n] - ws[n - 1]) / dt
            else:
                dwdt = (1.5 * ws[n] - 2.0 * ws[n - 1] + 0.5 * ws[n - 2]) / dt
            lhs = alpha * dwdt + beta * ws[n]
            assert np.abs(lhs - (mu2 * us[n - 1] - ws[n] ** 3)).max() < 1e-9, n
