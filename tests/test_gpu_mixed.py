"""Mixed precision (option "precision" = 1): FP32 storage of the condensed operator,
the local-solve factors and the preconditioner blocks, FP64 residuals and iterative
refinement.  The converged solution agrees with the FP64 oracle to 1e-8 relative
(the stated tolerance of the mixed form) and the final residual is at most
refine_rtol (1e-10) times the first one.
"""
import sys

import numpy as np
import pytest

from oracle import hdg as ora

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-8

CASES = {
    "cd_bj": ("convdiff3d", dict(n=2, porder=4, ncu=2, restart=100, gmres_tol=1e-12), "convdiff"),
    "cd_ilu": ("convdiff3d", dict(n=4, porder=2, ncu=2, restart=40, gmres_tol=1e-12, precond="bilu0",
                                  cartgridpart=[3, 4, 4, 4, 2, 2, 2, 0]), "convdiff"),
    "p2_eas": ("poisson2d", dict(n=8, porder=3, pgauss=8, precond=1), "poisson"),
    # Schwarz blocks inverted in FP32 and applied in BF16 (option precond_bf16): small blocks
    # (one thread per row), m = 192 (16-byte loads) and m = 300 (8-byte loads)
    "p2_eas_bf16": ("poisson2d", dict(n=8, porder=3, pgauss=8, precond=1), "poisson", {"precond_bf16": 1}),
    "cd3_eas_bf16": ("convdiff3d", dict(n=3, porder=3, ncu=2, restart=100, gmres_tol=1e-12, precond="eas",
                                        inflow_dirichlet=True), "convdiff", {"precond_bf16": 1}),
    "cd4_eas_bf16": ("convdiff3d", dict(n=2, porder=4, ncu=2, restart=100, gmres_tol=1e-12, precond="eas"),
                     "convdiff", {"precond_bf16": 1}),
    # the FP32 LU of D' (n = 250, 64-wide panels) and of the Schwarz blocks (m = 300) by substitution
    # instead of triangle inverses (option lu_inverses = 0)
    "cd_bj_subst": ("convdiff3d", dict(n=2, porder=4, ncu=2, restart=100, gmres_tol=1e-12), "convdiff",
                    {"lu_inverses": 0}),
    "cd4_eas_subst": ("convdiff3d", dict(n=2, porder=4, ncu=2, restart=100, gmres_tol=1e-12, precond="eas"),
                      "convdiff", {"lu_inverses": 0}),
    # each solve capped at one restart cycle of 10: the refinement loop carries the convergence
    "cd_ilu_cycles": ("convdiff3d", dict(n=4, porder=2, ncu=2, restart=10, gmres_tol=1e-12, precond="bilu0",
                                         cartgridpart=[3, 4, 4, 4, 2, 2, 2, 0]), "convdiff",
                      {"refine_cycles": 1, "refine_maxit": 60, "refine_first_tol": 1e-7, "refine_tol": 1e-7}),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def built(case_dir):
    out = {}

    def get(name):
        if name not in out:
            builder, kw, model = CASES[name][:3]
            path, info = case_dir("mixed_" + name, builder, **kw)
            out[name] = (path, model)
        return out[name]
    return get


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_solve_matches_fp64_oracle(built, name):
    path, model = built(name)
    say("case %s: mixed-precision solve" % name)
    from exasim_amd.hdg import HDGSolver
    P = ora.HDGProblem(path, model)
    r0 = None
    ora.newton_solve(P)
    with HDGSolver(path, model) as S:
        S.set_option("precision", 1)
        for k, v in (CASES[name][3] if len(CASES[name]) > 3 else {}).items():
            S.set_option(k, v)
        S.get_q()
        nb, nru = S.residual()
        r0 = nb + nru
        S.reset()
        it, gits, nrm = S.solve()
        z = S.sizes
        uh = S.get("uh")
        udg = S.get("udg").reshape(z["ne"], z["nc"], z["npe"]).transpose(0, 2, 1)
    say("  gmres %d, residual %.3e -> %.3e, err uh %.2e" % (gits, r0, nrm, rel(uh, P.uh.reshape(-1))))
    assert nrm <= 1e-10 * r0 * 1.0001 or nrm < 1e-14
    assert rel(uh, P.uh.reshape(-1)) < MIXED_TOL
    assert rel(udg, P.udg) < MIXED_TOL


def test_mixed_matvec_uses_fp32_operator(built):
    """The mixed matvec multiplies by the FP32 H (assembled in FP32 storage) with FP64
    accumulation; that H agrees with the FP64 assembly's to FP32 accuracy."""
    path, model = built("cd_bj")
    say("case cd_bj: FP32 matvec")
    from exasim_amd.hdg import HDGSolver
    with HDGSolver(path, model) as S:
        S.get_q()
        z = S.sizes
        shape = lambda a: a.reshape(z["ne"], z["m"], z["m"]).transpose(0, 2, 1)
        S.set_option("precision", 0)
        S.assemble()
        H64 = shape(S.get("H"))
        S.set_option("precision", 1)
        S.assemble()
        H = shape(S.get("H"))
        v = np.random.default_rng(9).standard_normal(z["N"])
        px, _ = S.device_ptr("x")
        pw, _ = S.device_ptr("w")
        S.set("x", v)
        S.matvec(px, pw)
        w = S.get("w")
    assert np.array_equal(H, H.astype(np.float32).astype(np.float64))   # FP32 values
    assert rel(H, H64) < 1e-4
    P = ora.HDGProblem(path, model)
    assert rel(w, P.matvec(H, v)) < 1e-12
    assert rel(w, P.matvec(H64, v)) > 1e-10          # it is not the FP64 operator


def test_bf16_ilu_factors_solve_to_tolerance(built):
    """precond_bf16: the mixed form with BF16 block-ILU(0) factors still reaches the FP64
    oracle's solution to 1e-8 (the preconditioner only changes the iterations)."""
    path, model = built("cd_ilu")
    say("case cd_ilu: mixed form with BF16 block-ILU(0) factors")
    from exasim_amd.hdg import HDGSolver
    P = ora.HDGProblem(path, model)
    ora.newton_solve(P)
    with HDGSolver(path, model) as S:
        S.set_option("precision", 1)
        S.set_option("precond_bf16", 1)
        S.solve()
        uh = S.get("uh")
    assert rel(uh, P.uh.reshape(-1)) < 1e-8


def test_bilu_factorisation_reproducible(case_dir):
    """Block-ILU(0) factorisation of 512 superelements with 50 x 50 face blocks (p = 4, two
    components): two factorisations give the same factors bit for bit, in FP64 and in the
    mixed form's FP32, and the FP32 factors agree with the FP64 ones block by block (a
    regression test for the Gauss-Jordan inverse of the diagonal blocks, whose pivot row is
    scaled only after every wave has read the pivot)."""
    path, _ = case_dir("cd16_ilu", "convdiff3d", n=16, porder=4, ncu=2, restart=100, gmres_tol=1e-10,
                       precond="bilu0", cartgridpart=[3, 16, 16, 16, 2, 2, 2, 0])
    say("case cd16_ilu: block-ILU(0) factors of 512 superelements, FP64 and FP32, twice each")
    from exasim_amd.hdg import HDGSolver
    with HDGSolver(path, "convdiff") as S:
        S.get_q()
        fac = {}
        for prec in (0, 1):
            S.set_option("precision", prec)
            S.assemble()
            a = S.get("BEilu")
            S.assemble()
            b = S.get("BEilu")
            assert np.isfinite(a).all()
            assert np.array_equal(a, b), prec
            fac[prec] = a.reshape(-1, 50 * 50)
    d = np.abs(fac[1] - fac[0]).max(axis=1) / np.maximum(np.abs(fac[0]).max(axis=1), 1e-300)
    assert d.max() < 1e-3, d.max()
