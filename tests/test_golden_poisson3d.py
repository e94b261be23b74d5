"""The reference's own committed run: examples/Poisson/poisson3d (tet P2,
24,576 elements, 304,128 trace unknowns), dataout/outuhat_np0.bin.

The reference solved it with GMRES(50) to a relative tolerance of 1e-8, so its
uhat is compared at 5e-8 relative.  Its boundary-face order comes from a hash
map; all boundary uhat values are zero (Dirichlet 0), so the sample holds
interior faces, whose order is fixed by the element order."""
import numpy as np
import pytest

from oracle import hdg as ora
from conftest import make_poisson3d_datain

TOL = 5e-8


@pytest.fixture(scope="module")
def p3d(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("poisson3d"))
    faces, ref, hdr = make_poisson3d_datain(path)
    return path, faces, ref, hdr


def test_oracle_reproduces_reference_uhat(p3d):
    path, faces, ref, hdr = p3d
    P = ora.HDGProblem(path, "poisson")
    assert P.preconditioner == 1          # the reference ran the elemental additive Schwarz
    assert (P.ncu, P.npf, P.nf) == tuple(int(v) for v in hdr)
    P.gmres_tol = 1e-10
    P.newton_maxit = 1
    ora.newton_solve(P)
    check(P.uh.reshape(P.nf, -1), P.conn["nif"], faces, ref)


def check(uh, nif, faces, ref):
    inner = faces < nif
    err = np.abs(uh[faces[inner]] - ref[inner]).max() / np.abs(ref).max()
    assert err < TOL, err
    # boundary faces: Dirichlet zero on both sides (their order differs, see module doc)
    assert np.abs(ref[~inner]).max() < 1e-12 and np.abs(uh[nif:]).max() < 1e-12


@pytest.mark.gpu
def test_gpu_reproduces_reference_uhat(p3d):
    from exasim_amd.hdg import HDGSolver
    path, faces, ref, hdr = p3d
    print("case poisson3d golden on GPU", flush=True)
    with HDGSolver(path, "poisson") as S:
        assert S.sizes["preconditioner"] == 1
        S.set_option("gmres_tol", 1e-10)
        S.set_option("newton_maxit", 1)
        S.solve()
        uh = S.get("uh").reshape(S.sizes["nf"], -1)
        nif = int((S.get("f2e").reshape(-1, 4)[:, 2] >= 0).sum())
    check(uh, nif, faces, ref)
