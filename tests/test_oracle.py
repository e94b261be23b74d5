"""Pins of the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign, a transposed operand, a wrong index map) fails it:

* master element and mesh nodes   -> the reference's own data tables and C1 xdg.bin
* connectivity                    -> both sides of every face see the same physical nodes
* geometry, mass, q-equation      -> closed forms (areas, outward normals, q = -grad u
                                     exact for linear u)
* assembly + condensation         -> HDG reproduces polynomials of degree <= p to round-off;
                                     dense block elimination of the local (u, q, uhat) system;
                                     finite differences of the residual
* GMRES                           -> dense direct solve
* full solve                      -> manufactured solution and its convergence rate
"""
import os

import numpy as np
import pytest

from exasim_amd import cases, datain as dw
from oracle import hdg as ora

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def tables():
    return np.load(os.path.join(GOLD, "text2code_tables.npz"))


@pytest.mark.parametrize("nd,p", [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 2), (3, 4)])
def test_master_nodes_and_perm_match_text2code_tables(tables, nd, p):
    m = dw.make_master(nd, p, 2 * p)
    assert np.array_equal(m["xpe"], tables["xpe_d%d_p%d" % (nd, p)][:, :nd])
    assert np.array_equal(m["xpf"], tables["xpf_d%d_p%d" % (nd, p)][:, :nd - 1])
    assert np.array_equal(m["perm"].T, tables["perm_d%d_p%d" % (nd, p)])


@pytest.mark.parametrize("nd,pg", [(2, 2), (2, 8), (2, 14), (2, 15), (3, 4), (3, 8)])
def test_gauss_rules_match_text2code_tables(tables, nd, pg):
    m = dw.make_master(nd, 2, pg)
    assert np.array_equal(m["gpe"], tables["gp_d%d_g%d" % (nd, pg)])
    assert np.abs(m["gwe"] - tables["gw_d%d_g%d" % (nd, pg)]).max() < 1e-17
    assert np.array_equal(m["gpf"], tables["gp_d%d_g%d" % (nd - 1, pg)])


def test_c1_mesh_nodes_match_reference_xdg():
    z = np.load(os.path.join(GOLD, "poisson2d_xdg.npz"))
    ref = z["xdg"].reshape(256, 2, 16).transpose(0, 2, 1)
    m = dw.make_master(2, 3, 14)
    p, t = dw.structured_grid([16, 16])
    assert np.array_equal(dw.compute_dgnodes(m, p, t), ref)


@pytest.mark.parametrize("nd,p", [(2, 3), (3, 2)])
def test_shape_functions(nd, p):
    """Partition of unity, exact gradients of linear functions, exact quadrature."""
    m = dw.make_master(nd, p, 2 * p)
    sh = m["shapeg"]
    assert np.abs(sh[0].sum(1) - 1).max() < 1e-13
    for a in range(nd):
        assert np.abs(sh[1 + a].sum(1)).max() < 1e-12
        grad = sh[1 + a] @ m["xpe"][:, a]
        assert np.abs(grad - 1).max() < 1e-12
    # int_0^1 x^(2p) over the cube/square = 1/(2p+1) per direction
    val = (sh[0] @ (m["xpe"][:, 0] ** p)) ** 2
    assert abs(val @ m["gwe"] - 1.0 / (2 * p + 1)) < 1e-13


@pytest.fixture(scope="module")
def c1small(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("c1s"))
    cases.poisson2d(path, n=3, porder=2, pgauss=6)
    return path


@pytest.fixture(scope="module")
def cd3(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("cd3"))
    cases.convdiff3d(path, n=2, porder=2, ncu=2, gmres_tol=1e-13)
    return path


@pytest.mark.parametrize("which", ["c1small", "cd3"])
def test_connectivity_faces_match_physically(which, c1small, cd3):
    P = ora.HDGProblem(c1small if which == "c1small" else cd3, "poisson" if which == "c1small" else "convdiff")
    fc = P.conn["facecon"]                       # [f, j, side] = e*npe + node
    x = P.xdg.reshape(-1, P.nd)
    assert np.abs(x[fc[:, :, 0]] - x[fc[:, :, 1]]).max() < 1e-14
    # elemcon: each element face node maps to the global face node at the same place
    for e in range(P.ne):
        for l in range(P.nfe):
            g = P.elemcon[e, l]
            f = g // P.npf
            side = 0 if P.conn["f2e"][f[0], 0] == e and P.conn["f2e"][f[0], 1] == l else 1
            assert np.abs(x[fc[f, g % P.npf, side]] - P.xdg[e, P.perm[l]]).max() < 1e-14
    nint = P.conn["nif"]
    assert np.all(P.conn["f2e"][:nint, 2] >= 0) and np.all(P.conn["f2e"][nint:, 2] == -1)


@pytest.mark.parametrize("which", ["c1small", "cd3"])
def test_geometry_closed_forms(which, c1small, cd3):
    P = ora.HDGProblem(c1small if which == "c1small" else cd3, "poisson" if which == "c1small" else "convdiff")
    vol = (P.jac * P.master["gwe"][None, :]).sum()
    assert abs(vol - 1.0) < 1e-13
    area = (P.jacf * P.master["gwf"][None, None, :]).sum(axis=2)     # [e, l]
    h = 1.0 / round(P.ne ** (1.0 / P.nd))
    assert np.abs(area - h ** (P.nd - 1)).max() < 1e-13
    cen = P.xdg.mean(axis=1)
    out = np.einsum("elgd,elgd->elg", P.nlf, P.xgf - cen[:, None, None, :])
    assert out.min() > 0
    assert np.abs(np.linalg.norm(P.nlf, axis=-1) - 1).max() < 1e-14


@pytest.mark.parametrize("which", ["c1small", "cd3"])
def test_q_equation_exact_for_linear_u(which, c1small, cd3):
    """q = -grad u exactly when u and uhat are traces of a linear function."""
    P = ora.HDGProblem(c1small if which == "c1small" else cd3, "poisson" if which == "c1small" else "convdiff")
    a = np.array([0.3, -1.7, 0.9])[:P.nd]
    udg = np.zeros_like(P.udg)
    udg[:, :, :P.ncu] = (P.xdg @ a)[:, :, None]
    uh = udg[P.conn["f2e"][:, 0][:, None], P.perm[P.conn["f2e"][:, 1]], :P.ncu]
    P.get_q(udg, uh)
    for c in range(P.ncu):
        for d in range(P.nd):
            assert np.abs(udg[:, :, P.ncu + c + P.ncu * d] + a[d]).max() < 1e-12
    assert np.abs(P.Mass - np.transpose(P.Mass, (0, 2, 1))).max() < 1e-16


@pytest.mark.parametrize("nd,porder,n", [(2, 2, 3), (2, 3, 2), (3, 2, 2)])
def test_polynomial_reproduction(tmp_path, nd, porder, n):
    cases.poisson2d(str(tmp_path), n=n, porder=porder, pgauss=2 * porder + 2, model="poisson_poly",
                    newton_iter=2, nd=nd)
    P = ora.HDGProblem(str(tmp_path), "poisson_poly")
    ora.newton_solve(P)
    x = P.xdg.reshape(-1, nd)
    assert np.abs(P.udg[:, :, 0].reshape(-1) - np.sum(x * x, 1)).max() < 1e-10
    for d in range(nd):
        assert np.abs(P.udg[:, :, 1 + d].reshape(-1) + 2 * x[:, d]).max() < 1e-9


def test_condensation_equals_dense_block_elimination(cd3):
    """H_cond = H - [K G] A^{-1} [F; E~], b_cond = Rh - [K G] A^{-1} [Ru; 0] with the
    local (u, q) block A = [[D, B], [-C~, M~]] assembled densely."""
    P = ora.HDGProblem(cd3, "convdiff")
    P.get_q(P.udg, P.uh)
    E = P.element_matrices()
    S = P.schur(E)
    npe, ncu, nd, n, m = P.npe, P.ncu, P.nd, P.n, P.m
    eye = np.eye(ncu)
    for e in range(P.ne):
        Bfull = np.concatenate([E["B"][d][e] for d in range(nd)], axis=1)          # n x n*nd
        Gfull = np.concatenate([E["G"][d][e] for d in range(nd)], axis=1)
        M = np.kron(eye, P.Mass[e])                                                 # (c, node) order
        Ct = np.concatenate([np.kron(eye, P.Mass[e] @ P.MinvC[d][e]) for d in range(nd)], axis=0)
        Et = np.concatenate([np.einsum("ab,kt->aktb", eye, P.Mass[e] @ P.MinvE[d][e]).reshape(n, m)
                             for d in range(nd)], axis=0)
        A = np.block([[E["D"][e], Bfull], [-Ct, np.kron(np.eye(nd), M)]])
        rhs_F = np.concatenate([E["F"][e], Et], axis=0)
        rhs_R = np.concatenate([E["Ru"][e], np.zeros(n * nd)])
        KG = np.concatenate([E["K"][e], Gfull], axis=1)
        Hc = E["H"][e] - KG @ np.linalg.solve(A, rhs_F)
        bc = E["Rh"][e] - KG @ np.linalg.solve(A, rhs_R)
        assert np.abs(Hc - S["H"][e]).max() < 1e-10 * np.abs(Hc).max()
        assert np.abs(bc - S["Rh"][e]).max() < 1e-10 * max(np.abs(bc).max(), 1e-300)


def test_jacobian_matches_finite_differences(cd3):
    """D, B, F, K, G, H are the derivatives of (-Ru, -Rh) w.r.t. (u, q, uhat)."""
    P = ora.HDGProblem(cd3, "convdiff")
    rng = np.random.default_rng(3)
    P.udg[:] = rng.standard_normal(P.udg.shape)
    P.uh[:] = rng.standard_normal(P.uh.shape)
    E = P.element_matrices()
    du = rng.standard_normal(P.udg.shape)
    duh = rng.standard_normal(P.uh.shape)
    eps = 1e-6

    def res(s):
        _, Ru, Rh = P.residual(P.udg + s * du, P.uh + s * duh, elem=True)
        return Ru, Rh

    (Rp, Hp), (Rm, Hm) = res(eps), res(-eps)
    fd = (Rp - Rm) / (2 * eps)
    fdh = (Hp - Hm) / (2 * eps)
    npe, ncu, nd = P.npe, P.ncu, P.nd
    dU = du[:, :, :ncu].transpose(0, 2, 1).reshape(P.ne, -1)
    dQ = [du[:, :, ncu + ncu * d:ncu + ncu * (d + 1)].transpose(0, 2, 1).reshape(P.ne, -1) for d in range(nd)]
    dUH = P.gather_uh(P.uh * 0 + duh).reshape(P.ne, -1)
    lin = -(np.einsum("eab,eb->ea", E["D"], dU) + sum(np.einsum("eab,eb->ea", E["B"][d], dQ[d]) for d in range(nd))
            + np.einsum("eab,eb->ea", E["F"], dUH))
    assert np.abs(fd - lin).max() < 1e-7 * np.abs(lin).max()
    linh = -(np.einsum("eab,eb->ea", E["K"], dU) + sum(np.einsum("eab,eb->ea", E["G"][d], dQ[d]) for d in range(nd))
             + np.einsum("eab,eb->ea", E["H"], dUH))
    assert np.abs(fdh - linh).max() < 1e-7 * np.abs(linh).max()


def test_gmres_matches_direct_solve(c1small):
    P = ora.HDGProblem(c1small, "poisson")
    P.get_q(P.udg, P.uh)
    E, S, b = P.assemble()
    N = b.size
    A = np.stack([P.matvec(S["H"], np.eye(N)[i]) for i in range(N)], 1)
    BE = P.block_jacobi(S["H"])
    x, its = ora.gmres(lambda v: P.matvec(S["H"], v), lambda r: P.apply_block_jacobi(BE, r), b, 50, 500,
                       1e-12, 1e-14)
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xd) < 1e-10 * np.linalg.norm(xd)
    # the block-Jacobi inverse blocks are the inverse diagonal face blocks of A
    ncf = P.ncu * P.npf
    for f in range(P.nf):
        blk = A[f * ncf:(f + 1) * ncf, f * ncf:(f + 1) * ncf]
        assert np.abs(np.linalg.inv(blk) - BE[f]).max() < 1e-10 * np.abs(BE[f]).max()


def test_c1_manufactured_solution_converges(tmp_path):
    errs = []
    for n in (4, 8):
        path = str(tmp_path / ("n%d" % n))
        cases.poisson2d(path, n=n, porder=3, pgauss=14)
        P = ora.HDGProblem(path, "poisson")
        ora.newton_solve(P)
        x = P.xdg.reshape(-1, 2)
        errs.append(np.abs(P.udg[:, :, 0].reshape(-1) - P.model.exact(x)).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > 3.5, (errs, rate)


def test_block_jacobi_and_schwarz_blocks_match_global_matrix(c1small):
    """Block-Jacobi blocks are the inverses of the global matrix's diagonal face blocks;
    the elemental additive Schwarz block of an element is its H_e with each face's
    diagonal block replaced by the global one (the global matrix is built column by
    column from the matvec)."""
    P = ora.HDGProblem(c1small, "poisson")
    P.get_q(P.udg, P.uh)
    E, S, b = P.assemble()
    H = S["H"]
    ncf = P.ncu * P.npf
    N = P.nf * ncf
    Ag = np.stack([P.matvec(H, e) for e in np.eye(N)], 1)
    BJ = P.block_jacobi(H)
    for f in range(P.nf):
        dof = f * ncf + np.arange(ncf)
        assert np.abs(np.linalg.inv(BJ[f]) - Ag[np.ix_(dof, dof)]).max() < 1e-11 * np.abs(Ag).max()
    BEs = np.linalg.inv(P.elemental_schwarz(H))
    for e in range(P.ne):
        dof = (np.arange(P.ncu)[None, :] + P.ncu * P.elemcon[e].reshape(-1)[:, None]).reshape(-1)
        ref = H[e].copy()
        for l in range(P.nfe):
            loc = slice(l * ncf, (l + 1) * ncf)
            ref[loc, loc] = Ag[np.ix_(dof[loc], dof[loc])]
        assert np.abs(BEs[e] - ref).max() < 1e-10 * np.abs(ref).max()
