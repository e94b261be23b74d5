"""GPU parity of the models generated from pdemodel files (exasim_amd/models/ through
modelgen into libexahip.so).

* poisson2d / convdiff3d restate the built-in poisson / convdiff models: the generated
  code (dual-number Jacobians) gives the same assembled system and solution as the
  hand-written one, to 1e-12 relative.
* ns2d (compressible Navier-Stokes, isothermal wall, characteristic far field) against
  the oracle's interpreter of the same file: the assembled system at the initial state
  to 1e-10, and the state after DIRK(3,3) steps to 1e-10 relative, with block-Jacobi;
  with block-ILU(0) the converged state agrees to 1e-9, and the mixed-precision form
  (FP32 Jacobian and preconditioner, FP64 residual and refinement) to 1e-8.
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
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def as_elem_mats(flat, rows, cols, ne):
    return flat.reshape(ne, cols, rows).transpose(0, 2, 1)


def solver(path, model):
    from exasim_amd.hdg import HDGSolver
    return HDGSolver(path, model)


def system(S):
    """(H, DinvF, Ru, Rh, b, udg) after get_q + assemble, as [e, ...] arrays."""
    z = S.sizes
    ne, n, m = z["ne"], z["n"], z["m"]
    S.get_q()
    S.assemble()
    return dict(H=as_elem_mats(S.get("H"), m, m, ne), DinvF=as_elem_mats(S.get("DinvF"), n, m, ne),
                Ru=S.get("Ru").reshape(ne, n), Rh=S.get("Rh").reshape(ne, m), b=S.get("b"),
                udg=S.get("udg").reshape(ne, z["nc"], z["npe"]).transpose(0, 2, 1))


TWINS = {
    "poisson2d": ("poisson2d", dict(n=4, porder=3, pgauss=8), "poisson"),
    "convdiff3d": ("convdiff3d", dict(n=2, porder=3, ncu=2, restart=100, gmres_tol=1e-12), "convdiff"),
}


@pytest.mark.parametrize("gen", list(TWINS))
def test_generated_model_matches_builtin(case_dir, gen):
    builder, kw, builtin = TWINS[gen]
    say("generated model %s vs built-in %s" % (gen, builtin))
    path, _ = case_dir("twin_" + gen, builder, **kw)
    out = {}
    for model in (builtin, gen):
        with solver(path, model) as S:
            sysm = system(S)
            S.solve()
            sysm["uh"] = S.get("uh")
            sysm["udg_sol"] = S.get("udg")
            out[model] = sysm
    a, b = out[gen], out[builtin]
    for k in ("H", "DinvF", "Ru", "Rh", "b", "uh", "udg_sol"):
        assert rel(a[k], b[k]) < 1e-12, k


NS_SMALL = dict(nx=8, ny=2, porder=3, length=(2.5, 1.0), nsteps=1, precond="bjacobi", restart=200,
                gmres_tol=1e-12, newton_tol=1e-10)
NS_DIRK = dict(nx=16, ny=4, porder=4, length=(5.0, 1.0), nsteps=2, precond="bjacobi", restart=200,
               gmres_tol=1e-12, newton_tol=1e-10)


def test_ns_assembly_matches_oracle(case_dir):
    say("ns2d: assembly at the initial state (8x2, p=3)")
    path, _ = case_dir("ns_small", "navierstokes2d", **NS_SMALL)
    P = ora.HDGProblem(path, "ns2d")
    P.get_q(P.udg, P.uh)
    E, S_, b = P.assemble()
    with solver(path, "ns2d") as S:
        g = system(S)
    assert rel(g["udg"], P.udg) < 1e-10
    for k in ("H", "DinvF", "Ru", "Rh"):
        assert rel(g[k], S_[k]) < 1e-10, k
    assert rel(g["b"], b) < 1e-10


@pytest.fixture(scope="module")
def ns_oracle(case_dir):
    path, _ = case_dir("ns_dirk", "navierstokes2d", **NS_DIRK)
    P = ora.HDGProblem(path, "ns2d")
    ora.dirk_solve(P)
    return path, P


def _state(S):
    z = S.sizes
    return S.get("uh"), S.get("udg").reshape(z["ne"], z["nc"], z["npe"]).transpose(0, 2, 1)


def test_ns_dirk_matches_oracle(ns_oracle):
    path, P = ns_oracle
    say("ns2d: DIRK(3,3) 2 steps (16x4, p=4), block-Jacobi, vs oracle")
    with solver(path, "ns2d") as S:
        S.solve()
        uh, udg = _state(S)
    assert rel(uh, P.uh.reshape(-1)) < 1e-10
    assert rel(udg[:, :, :4], P.udg[:, :, :4]) < 1e-10


def test_ns_bilu0_and_mixed_precision(ns_oracle):
    path, P = ns_oracle
    say("ns2d: block-ILU(0) FP64 and mixed precision vs oracle")
    with solver(path, "ns2d") as S:
        S.set_option("preconditioner", 2)
        S.solve()
        uh64, udg64 = _state(S)
        S.reset()
        S.set_option("precision", 1)
        S.solve()
        uh32, udg32 = _state(S)
    assert rel(uh64, P.uh.reshape(-1)) < 1e-9
    assert rel(udg64[:, :, :4], P.udg[:, :, :4]) < 1e-9
    assert rel(uh32, P.uh.reshape(-1)) < 1e-8
    assert rel(udg32[:, :, :4], P.udg[:, :, :4]) < 1e-8
