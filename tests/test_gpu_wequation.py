"""GPU parity of the w equation (models with a Sourcew function; wequation.cpp:31-240) against
the oracle: the point Newton solve for w in the volume and face kernels, the chain rule through
w_uq, the nodal update of w (GetW) and its time-stepping source under DIRK and BDF.

* wsub2d (w = u) assembles the system of wsub2d_ref (w written out) to 1e-12 on the GPU;
* steady and time-dependent wnl2d solutions (u, uhat and the nodal w) agree with the oracle to
  1e-10 relative;
* a batch of two wnl2d instances gives each instance's result alone, bit for bit.
"""
import sys

import numpy as np
import pytest

from oracle import hdg as ora

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def test_w_chain_rule_on_gpu(tmp_path):
    say("wsub2d against wsub2d_ref: assembled system")
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver
    a, b = str(tmp_path / "w"), str(tmp_path / "ref")
    cases.wmodel2d(a, model="wsub2d", mu=(0.5, 1.0), u0=0.3, w0=0.0)
    cases.wmodel2d(b, model="wsub2d_ref", mu=(0.5, 1.0), u0=0.3)
    out = {}
    for name, path in (("wsub2d", a), ("wsub2d_ref", b)):
        with HDGSolver(path, name) as S:
            S.get_q()
            S.assemble()
            out[name] = {k: S.get(k) for k in ("H", "DinvF", "Ru", "Rh", "b")}
    for k in out["wsub2d_ref"]:
        assert rel(out["wsub2d"][k], out["wsub2d_ref"][k]) < 1e-12, k


CASES = {
    "steady": dict(mu=(1.0, 1.0, 2.0), alpha=0.0, beta=1.0, u0=0.0, w0=0.0),
    "dirk22": dict(mu=(1.0, 1.0, 2.0), alpha=1.0, beta=1.0, u0=0.0, w0=0.2, dt=[0.05] * 3, torder=2,
                   nstage=2, scheme=0),
    "bdf2": dict(mu=(1.0, 1.0, 2.0), alpha=1.0, beta=1.0, u0=0.0, w0=0.2, dt=[0.05] * 3, torder=2,
                 scheme=1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_w_solve_matches_oracle(tmp_path, name):
    say("wnl2d %s: %s" % (name, CASES[name]))
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver
    path = str(tmp_path / name)
    cases.wmodel2d(path, model="wnl2d", **CASES[name])
    P = ora.HDGProblem(path, "wnl2d")
    if P.tdep:
        ora.time_solve(P)
    else:
        ora.newton_solve(P)
    with HDGSolver(path, "wnl2d") as S:
        S.solve()
        z = S.sizes
        uh = S.get("uh")
        udg = S.get("udg").reshape(z["ne"], z["nc"], z["npe"]).transpose(0, 2, 1)
        wdg = S.get("wdg").reshape(z["ne"], 1, z["npe"]).transpose(0, 2, 1)
    assert rel(uh, P.uh.reshape(-1)) < 1e-10
    assert rel(udg, P.udg) < 1e-10
    assert rel(wdg, P.wdg) < 1e-10


def test_w_batch_equals_alone(tmp_path):
    say("wnl2d batch of two (different mu) against each alone")
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver
    paths = []
    for i, mu in enumerate([(1.0, 1.0, 2.0), (0.5, 2.0, 1.0)]):
        p = str(tmp_path / ("i%d" % i))
        cases.wmodel2d(p, model="wnl2d", mu=mu, alpha=1.0, beta=1.0, w0=0.2, dt=[0.05] * 2, torder=1, nstage=1)
        paths.append(p)
    with HDGSolver.batch(paths, "wnl2d") as S:
        S.solve()
        together = [S.get_instance(i, "uh") for i in range(2)]
    for i, p in enumerate(paths):
        with HDGSolver.batch([p], "wnl2d") as S:
            S.solve()
            assert np.array_equal(S.get_instance(0, "uh"), together[i]), i
