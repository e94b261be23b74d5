"""GPU parity of the superelement block-ILU(0) preconditioner (preconditioner = 2)
against the oracle (oracle/bilu0.py): face ordering and CRS pattern bit-exact,
factored blocks and the preconditioner apply to 1e-10 relative, and the
preconditioned GMRES solve (modified and classical Gram-Schmidt) to 1e-10 relative with the
same iteration count.
"""
import sys

import numpy as np
import pytest

from oracle import hdg as ora
from oracle import bilu0 as obilu

pytestmark = pytest.mark.gpu

RTOL = 1e-10

CASES = {
    # 3D convection-diffusion, 4^3 hexes in 2x2x2 superelements, p = 2
    "cd3": ("convdiff3d", dict(n=4, porder=2, ncu=2, restart=40, gmres_tol=1e-12, precond="bilu0",
                               cartgridpart=[3, 4, 4, 4, 2, 2, 2, 0]), "convdiff"),
    # same mesh, superelements enumerated with the other loop order (opt = 1), 4x2x4 blocks
    "cd3o": ("convdiff3d", dict(n=(4, 4, 4), porder=1, ncu=1, restart=40, gmres_tol=1e-12,
                                precond="bilu0", cartgridpart=[3, 4, 4, 4, 4, 2, 4, 1]), "convdiff"),
    # 2D Poisson, 8x8 quads in 4x4 superelements, p = 3
    "p2": ("poisson2d", dict(n=8, porder=3, pgauss=8, restart=30, gmres_tol=1e-12, precond="bilu0",
                             cartgridpart=[2, 8, 8, 4, 4, 0]), "poisson"),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def built(case_dir):
    out = {}

    def get(name):
        if name not in out:
            builder, kw, model = CASES[name]
            path, info = case_dir(name, builder, **kw)
            out[name] = (path, model)
        return out[name]
    return get


def oracle_structure(P):
    elem = obilu.gridpartition(P.mesh["cartgridpart"])
    row_ptr, col_ind, face = obilu.crs_faceordering(elem, P.conn["f2e"], P.nfe)
    return row_ptr, col_ind, face


@pytest.mark.parametrize("name", list(CASES))
def test_bilu0_structure_bit_exact(built, name):
    path, model = built(name)
    say("case %s: block-ILU(0) structure" % name)
    from exasim_amd.hdg import HDGSolver
    P = ora.HDGProblem(path, model)
    row_ptr, col_ind, face = oracle_structure(P)
    with HDGSolver(path, model) as S:
        gface = S.get("ilu_face").astype(np.int64)
        grp = S.get("ilu_row_ptr").astype(np.int64)
        gci = S.get("ilu_col_ind").astype(np.int64)
    nse, nfse = face.shape
    assert np.array_equal(grp, row_ptr)
    assert np.array_equal(gci, col_ind)
    assert np.array_equal(gface.reshape(nfse, nse).T, face)


@pytest.mark.parametrize("name", list(CASES))
def test_bilu0_factor_and_apply(built, name):
    path, model = built(name)
    say("case %s: block-ILU(0) factor/apply" % name)
    from exasim_amd.hdg import HDGSolver
    P = ora.HDGProblem(path, model)
    P.get_q(P.udg, P.uh)
    E, S_, b = P.assemble()
    row_ptr, col_ind, face = oracle_structure(P)
    ind_ii, elim, L, U = obilu.crs_indexingilu0(row_ptr, col_ind, face.shape[1])
    BE = obilu.assemble(S_["H"], P.conn["f2e"], P.elemcon, face, row_ptr, col_ind, P.npf, P.nfe, P.ncu)
    obilu.factor(BE, ind_ii, elim)
    ncf = P.ncu * P.npf
    v = np.random.default_rng(11).standard_normal(b.size)
    with HDGSolver(path, model) as S:
        S.get_q()
        S.assemble()
        BEg = S.get("BEilu").reshape(BE.shape[0], BE.shape[1], ncf, ncf).transpose(0, 1, 3, 2)
        pw, _ = S.device_ptr("w")
        S.set("w", v)
        S.precond(pw)
        pv = S.get("w")
    assert rel(BEg, BE) < RTOL
    assert rel(pv, obilu.apply(BE, ind_ii, L, U, face, v, ncf, P.nf)) < RTOL


@pytest.mark.parametrize("ortho", [0, 1])
@pytest.mark.parametrize("name", list(CASES))
def test_bilu0_solve_matches_oracle(built, name, ortho):
    path, model = built(name)
    say("case %s: block-ILU(0) preconditioned solve, %s" % (name, "CGS" if ortho else "MGS"))
    from exasim_amd.hdg import HDGSolver
    P = ora.HDGProblem(path, model)
    P.ortho = ortho
    P.get_q(P.udg, P.uh)
    E, S_, b = P.assemble()
    prec = P.make_preconditioner(S_["H"])
    xo, its_o = ora.gmres(lambda q: P.matvec(S_["H"], q), prec, b, P.restart, P.gmres_maxit,
                          P.gmres_tol, P.newton_tol, P.ortho)
    with HDGSolver(path, model) as S:
        S.set_option("ortho", ortho)
        S.get_q()
        S.assemble()
        pb, _ = S.device_ptr("b")
        px, _ = S.device_ptr("x")
        its = S.gmres(pb, px)
        xg = S.get("x")
    say("  gmres iterations: gpu %d oracle %d" % (its, its_o))
    assert its == its_o
    assert rel(xg, xo) < 1e-9
