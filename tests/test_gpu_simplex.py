"""GPU parity on triangles and tetrahedra written by the datain writer (make_master with
elemtype 0, simplex_grid): the solution, the assembled system and the GMRES count agree with the
oracle (1e-10 relative, counts equal)."""
import sys

import numpy as np
import pytest

from oracle import hdg as ora

pytestmark = pytest.mark.gpu

CASES = {
    "tri_p3": dict(nd=2, n=4, porder=3, pgauss=6, elemtype=0),
    "tet_p2": dict(nd=3, n=2, porder=2, pgauss=4, elemtype=0),
    "tet_p3_eas": dict(nd=3, n=2, porder=3, pgauss=6, elemtype=0, precond=1),
}


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_simplex_solve_matches_oracle(case_dir, name):
    print("case %s: %s" % (name, CASES[name]), flush=True)
    sys.stdout.flush()
    path, info = case_dir("simplex_" + name, "poisson2d", **CASES[name])
    from exasim_amd.hdg import HDGSolver
    P = ora.HDGProblem(path, "poisson")
    hist = ora.newton_solve(P)
    with HDGSolver(path, "poisson") as S:
        it, gits, nrm = S.solve()
        uh = S.get("uh")
        z = S.sizes
        udg = S.get("udg").reshape(z["ne"], z["nc"], z["npe"]).transpose(0, 2, 1)
    assert rel(uh, P.uh.reshape(-1)) < 1e-10
    assert rel(udg, P.udg) < 1e-10
    assert gits == sum(hist["gmres_iters"])
