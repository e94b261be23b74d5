"""GPU parity of BDF time stepping and of the Tdfunc time-derivative coefficient against the
oracle: the state after all time steps agrees to 1e-10 relative, and a reset replays the same
steps bit for bit."""
import sys

import numpy as np
import pytest

from oracle import hdg as ora

pytestmark = pytest.mark.gpu

RTOL = 1e-10

CASES = {
    "heat2d_bdf1": dict(nd=2, n=4, porder=3, dt=0.02, nsteps=3, torder=1, scheme=1),
    "heat2d_bdf2": dict(nd=2, n=4, porder=3, dt=0.02, nsteps=4, torder=2, scheme=1),
    "heat2d_bdf3": dict(nd=2, n=4, porder=3, dt=0.02, nsteps=5, torder=3, scheme=1),
    "cd3d_bdf2": dict(nd=3, n=2, porder=2, ncu=2, kappa=0.1, vel=(1.0, 0.5, 0.25), dt=0.05, nsteps=3,
                      torder=2, scheme=1),
    # (c0 + c1 x) u_t = kappa lap u through the model's Tdfunc, with both time schemes
    "tdfunc_dirk22": dict(nd=2, n=4, porder=3, kappa=0.5, dt=0.02, nsteps=3, torder=2, nstage=2,
                          capacity=(2.0, 1.0)),
    "tdfunc_bdf3": dict(nd=2, n=4, porder=3, kappa=0.5, dt=0.02, nsteps=4, torder=3, scheme=1,
                        capacity=(2.0, 1.0)),
}


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_time_scheme_matches_oracle(case_dir, name):
    print("case %s: time stepping %s" % (name, CASES[name]), flush=True)
    sys.stdout.flush()
    path, info = case_dir("ts_" + name, "convdiff_unsteady", **CASES[name])
    model = info["model"]
    from exasim_amd.hdg import HDGSolver
    P = ora.HDGProblem(path, model)
    ora.time_solve(P)
    with HDGSolver(path, model) as S:
        S.solve()
        z = S.sizes
        uh = S.get("uh")
        udg = S.get("udg").reshape(z["ne"], z["nc"], z["npe"]).transpose(0, 2, 1)
        S.reset()
        S.solve()                                    # the reset restores the initial time and history
        uh2 = S.get("uh")
    assert rel(uh, P.uh.reshape(-1)) < RTOL
    assert rel(udg[:, :, :P.ncu], P.udg[:, :, :P.ncu]) < RTOL
    assert np.array_equal(uh, uh2)
