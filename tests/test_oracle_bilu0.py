"""Pins of the block-ILU(0) oracle (oracle/bilu0.py), CPU only.

* grid partition       -> every element in exactly one superelement, each superelement an
                          axis-aligned M-block of the lexicographic grid
* face ordering / CRS  -> each superelement's faces are exactly the faces of its elements; the
                          CRS pattern is face adjacency (two faces share an element of the
                          superelement), with the diagonal first
* block assembly       -> the blocks on the pattern equal the global condensed matrix (built
                          column by column from the matvec) restricted to the superelement
* factorisation        -> the defining property of ILU(0): (L U)_ij = A_ij on the pattern,
                          with L unit lower and U upper triangular; on a superelement of one
                          element the pattern is dense and ILU(0) is the exact LU
* apply                -> forward/backward sweeps solve (L U) y = b per superelement and the
                          scatter adds both superelements' values on shared faces
"""
import numpy as np
import pytest

from exasim_amd import cases
from oracle import bilu0 as ob
from oracle import hdg as ora


@pytest.mark.parametrize("cart", [[3, 4, 4, 4, 2, 2, 2, 0], [3, 4, 2, 6, 2, 1, 3, 0], [3, 4, 4, 4, 4, 2, 4, 1],
                                  [2, 8, 4, 4, 2, 0], [2, 8, 4, 2, 4, 1]])
def test_gridpartition_blocks(cart):
    nd = cart[0]
    N, M = cart[1:1 + nd], cart[1 + nd:1 + 2 * nd]
    opt = cart[1 + 2 * nd]
    elem = ob.gridpartition(cart)
    ne = int(np.prod(N))
    assert elem.shape == (ne // int(np.prod(M)), int(np.prod(M)))
    assert np.array_equal(np.sort(elem.ravel()), np.arange(ne))
    for blk in elem:
        # grid coordinates of each element (x fastest when opt == 0, reversed when opt == 1)
        if opt == 0:
            c = np.array(np.unravel_index(blk, N[::-1]))[::-1]
        else:
            c = np.array(np.unravel_index(blk, N))
        for d in range(nd):
            assert c[d].max() - c[d].min() == M[d] - 1
            assert c[d].min() % M[d] == 0


def _problem(tmp_path, n, se, porder=1, nd=3):
    path = str(tmp_path)
    if nd == 3:
        cases.convdiff3d(path, n=n, porder=porder, ncu=1, precond="bilu0",
                         cartgridpart=[3, n, n, n, se, se, se, 0])
        P = ora.HDGProblem(path, "convdiff")
    else:
        cases.poisson2d(path, n=n, porder=porder, pgauss=2 * porder, precond="bilu0",
                        cartgridpart=[2, n, n, se, se, 0])
        P = ora.HDGProblem(path, "poisson")
    P.get_q(P.udg, P.uh)
    E, S, b = P.assemble()
    return P, S["H"]


@pytest.mark.parametrize("nd,n,se", [(3, 4, 2), (2, 6, 3), (3, 2, 1)])
def test_face_ordering_and_pattern(tmp_path, nd, n, se):
    P, H = _problem(tmp_path, n, se, nd=nd)
    f2e = P.conn["f2e"]
    elem = ob.gridpartition(P.mesh["cartgridpart"])
    row_ptr, col_ind, face = ob.crs_faceordering(elem, f2e, P.nfe)
    e2f = np.full((P.ne, P.nfe), -1)
    for f, (e1, l1, e2, l2) in enumerate(f2e):
        e2f[e1, l1] = f
        if e2 >= 0:
            e2f[e2, l2] = f
    for s in range(elem.shape[0]):
        fs = face[s]
        assert sorted(fs) == sorted(set(e2f[elem[s]].ravel()))     # exactly the faces, once each
        pos = {f: i for i, f in enumerate(fs)}
        adj = [set() for _ in fs]
        for e in elem[s]:
            loc = [pos[f] for f in e2f[e]]
            for a in loc:
                adj[a].update(loc)
        for i in range(len(fs)):
            cols = list(col_ind[row_ptr[i]:row_ptr[i + 1]])
            assert cols[0] == i and cols[1:] == sorted(cols[1:])
            assert set(cols) == adj[i]


def _dense_blocks(BE, s, row_ptr, col_ind, nfse, ncf):
    A = np.zeros((nfse * ncf, nfse * ncf))
    for i in range(nfse):
        for t in range(row_ptr[i], row_ptr[i + 1]):
            j = col_ind[t]
            A[i * ncf:(i + 1) * ncf, j * ncf:(j + 1) * ncf] = BE[t, s]
    return A


def _lu_factors(F, mask, ncf):
    """Unit-lower L and upper U from the factored blocks (diagonal blocks hold U_ii^{-1})."""
    nfse = mask.shape[0]
    Lm = np.eye(nfse * ncf)
    Um = np.zeros_like(Lm)
    for i in range(nfse):
        for j in range(nfse):
            if not mask[i, j]:
                continue
            blk = F[i * ncf:(i + 1) * ncf, j * ncf:(j + 1) * ncf]
            if j < i:
                Lm[i * ncf:(i + 1) * ncf, j * ncf:(j + 1) * ncf] = blk
            elif j == i:
                Um[i * ncf:(i + 1) * ncf, j * ncf:(j + 1) * ncf] = np.linalg.inv(blk)
            else:
                Um[i * ncf:(i + 1) * ncf, j * ncf:(j + 1) * ncf] = blk
    return Lm, Um


@pytest.mark.parametrize("nd,n,se", [(3, 4, 2), (2, 6, 3), (3, 2, 1)])
def test_ilu0_reproduces_matrix_on_pattern_and_apply_solves(tmp_path, nd, n, se):
    P, H = _problem(tmp_path, n, se, nd=nd)
    elem = ob.gridpartition(P.mesh["cartgridpart"])
    row_ptr, col_ind, face = ob.crs_faceordering(elem, P.conn["f2e"], P.nfe)
    nse, nfse = face.shape
    ncf = P.ncu * P.npf
    ind_ii, elim, L, U = ob.crs_indexingilu0(row_ptr, col_ind, nfse)
    A0 = ob.assemble(H, P.conn["f2e"], P.elemcon, face, row_ptr, col_ind, P.npf, P.nfe, P.ncu)
    BE = ob.factor(A0.copy(), ind_ii, elim)
    mask = np.zeros((nfse, nfse), dtype=bool)
    for i in range(nfse):
        mask[i, col_ind[row_ptr[i]:row_ptr[i + 1]]] = True
    big = np.kron(mask, np.ones((ncf, ncf), dtype=bool))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(P.nf * ncf)
    y = ob.apply(BE, ind_ii, L, U, face, x, ncf, P.nf).reshape(P.nf, ncf)
    xs = x.reshape(P.nf, ncf)
    acc = np.zeros_like(xs)
    # the assembled blocks are the restriction of the global condensed matrix
    # (columns from the matvec) to each superelement's faces, on the pattern
    N = P.nf * ncf
    Ag = np.stack([P.matvec(H, e) for e in np.eye(N)], 1)
    for s in range(nse):
        dof = (face[s][:, None] * ncf + np.arange(ncf)[None, :]).reshape(-1)
        A = _dense_blocks(A0, s, row_ptr, col_ind, nfse, ncf)
        assert np.abs((A - Ag[np.ix_(dof, dof)])[big]).max() < 1e-12 * np.abs(Ag).max()
    for s in range(nse):
        A = _dense_blocks(A0, s, row_ptr, col_ind, nfse, ncf)
        Lm, Um = _lu_factors(_dense_blocks(BE, s, row_ptr, col_ind, nfse, ncf), mask, ncf)
        LU = Lm @ Um
        assert np.abs((LU - A)[big]).max() < 1e-10 * np.abs(A).max()
        if se == 1:                       # one element: dense pattern, exact LU
            assert mask.all()
            assert np.abs(LU - A).max() < 1e-10 * np.abs(A).max()
        ys = np.linalg.solve(LU, xs[face[s]].reshape(-1)).reshape(nfse, ncf)
        np.add.at(acc, face[s], ys)
    assert np.abs(y - acc).max() < 1e-9 * np.abs(acc).max()
