"""GPU parity of the 3D convection-diffusion configuration the benchmark times (32^3 hexes,
p = 4, two components, GMRES(100) with classical Gram-Schmidt, block-ILU(0) on 2 x 2 x 2
superelements) at a size the oracle finishes in seconds: 4^3 hexes with the same p, ncu,
superelement shape, Krylov settings and face-block size (ncf = 50).

* FP64: the condensed system (H, D^-1 F, Ru, Rh, b), the factored block-ILU(0) blocks, the
  preconditioner apply and the preconditioned GMRES solve against the oracle (1e-10 relative;
  the GMRES iteration count equal), and the full Newton solve (1e-10).
* Mixed form with BF16 block-ILU(0) factors and the benchmark's refinement tolerances: the
  converged solution against the FP64 oracle at 1e-8.
* Assembly in several element blocks (a workspace of a few elements, the last block partial):
  every assembled array equals the one-block assembly (1e-13) in FP64 and in the mixed form,
  with each preconditioner, and the solve still matches the oracle.
"""
import sys

import numpy as np
import pytest

from oracle import hdg as ora
from oracle import bilu0 as obilu

pytestmark = pytest.mark.gpu

RTOL = 1e-10
KW = dict(n=4, porder=4, ncu=2, restart=100, gmres_tol=1e-10, precond="bilu0",
          cartgridpart=[3, 4, 4, 4, 2, 2, 2, 0], ortho=1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def mats(flat, rows, cols, ne):
    return flat.reshape(ne, cols, rows).transpose(0, 2, 1)


@pytest.fixture(scope="module")
def c2cfg(case_dir):
    path, _ = case_dir("c2cfg", "convdiff3d", **KW)
    say("case c2cfg (4^3 hex, p=4, ncu=2, CGS, block-ILU(0) 2x2x2): oracle assembly")
    P = ora.HDGProblem(path, "convdiff")
    assert P.ortho == 1
    P.get_q(P.udg, P.uh)
    E, S_, b = P.assemble()
    elem = obilu.gridpartition(P.mesh["cartgridpart"])
    row_ptr, col_ind, face = obilu.crs_faceordering(elem, P.conn["f2e"], P.nfe)
    ind_ii, elim, L, U = obilu.crs_indexingilu0(row_ptr, col_ind, face.shape[1])
    BE = obilu.assemble(S_["H"], P.conn["f2e"], P.elemcon, face, row_ptr, col_ind, P.npf, P.nfe, P.ncu)
    obilu.factor(BE, ind_ii, elim)
    ncf = P.ncu * P.npf
    assert ncf == 50
    v = np.random.default_rng(17).standard_normal(b.size)
    pv = obilu.apply(BE, ind_ii, L, U, face, v, ncf, P.nf)
    xo, its_o = ora.gmres(lambda q: P.matvec(S_["H"], q), P.make_preconditioner(S_["H"]), b, P.restart,
                          P.gmres_maxit, P.gmres_tol, P.newton_tol, P.ortho)
    say("  oracle GMRES (CGS) iterations %d; Newton solve" % its_o)
    Q = ora.HDGProblem(path, "convdiff")
    ora.newton_solve(Q)
    return dict(path=path, P=P, S=S_, b=b, BE=BE, v=v, pv=pv, xo=xo, its=its_o, uh=Q.uh.reshape(-1),
                udg=Q.udg, ilu=(ind_ii, L, U, face), ncf=ncf)


def test_c2cfg_fp64_stages(c2cfg):
    from exasim_amd.hdg import HDGSolver
    say("case c2cfg: FP64 assembly, block-ILU(0) factors and apply, GMRES")
    o = c2cfg
    with HDGSolver(o["path"], "convdiff") as S:
        z = S.sizes
        ne, n, m = z["ne"], z["n"], z["m"]
        assert (n, m, z["preconditioner"]) == (250, 300, 2)
        S.get_q()
        S.assemble()
        H = mats(S.get("H"), m, m, ne)
        DinvF = mats(S.get("DinvF"), n, m, ne)
        Ru = S.get("Ru").reshape(ne, n)
        Rh = S.get("Rh").reshape(ne, m)
        bg = S.get("b")
        BEg = S.get("BEilu").reshape(o["BE"].shape[0], o["BE"].shape[1], 50, 50).transpose(0, 1, 3, 2)
        pw, _ = S.device_ptr("w")
        S.set("w", o["v"])
        S.precond(pw)
        pv = S.get("w")
        pb, _ = S.device_ptr("b")
        px, _ = S.device_ptr("x")
        its = S.gmres(pb, px)
        xg = S.get("x")
    assert rel(H, o["S"]["H"]) < RTOL
    assert rel(DinvF, o["S"]["DinvF"]) < RTOL
    assert rel(Ru, o["S"]["Ru"]) < RTOL
    assert rel(Rh, o["S"]["Rh"]) < RTOL
    assert rel(bg, o["b"]) < RTOL
    assert rel(BEg, o["BE"]) < RTOL
    assert rel(pv, o["pv"]) < RTOL
    say("  gmres iterations: gpu %d oracle %d" % (its, o["its"]))
    assert its == o["its"]
    assert rel(xg, o["xo"]) < 1e-9


def bf16_round(a):
    """Round to BF16 (nearest, ties to even) as the library's FP32 -> BF16 conversion does."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("bf16", [0, 1])
def test_c2cfg_mixed_ilu_apply(c2cfg, bf16):
    """The block-ILU(0) apply of the mixed form's FP32 / BF16 factors, with one wave per
    superelement and with the 256-thread kernel, against the oracle's sweeps (ApplyBlockILU0,
    preconditioner.cpp:184-270) in FP64 on the same rounded factors. The kernels keep the
    superelement's vector in FP32 between row steps, hence the 1e-5 bound."""
    from exasim_amd.hdg import HDGSolver
    say("case c2cfg: mixed form block-ILU(0) apply, %s factors, wave and 256-thread kernels"
        % ("BF16" if bf16 else "FP32"))
    o = c2cfg
    with HDGSolver(o["path"], "convdiff") as S:
        S.set_option("precision", 1)
        S.set_option("precond_bf16", bf16)
        S.get_q()
        S.assemble()
        BEg = S.get("BEilu").reshape(o["BE"].shape[0], o["BE"].shape[1], 50, 50).transpose(0, 1, 3, 2)
        pw, _ = S.device_ptr("w")
        out = {}
        for wave in (1, 0):
            S.set_option("ilu_wave", wave)
            S.set("w", o["v"])
            S.precond(pw)
            out[wave] = S.get("w")
        S.set_option("ilu_wave", -1)
    # the FP32 factors are the FP32 factorisation of the FP32 condensed matrices: close to FP64's
    assert rel(BEg, o["BE"]) < 1e-3
    BEr = bf16_round(BEg) if bf16 else BEg.copy()
    ind_ii, L, U, face = o["ilu"]
    pv = obilu.apply(BEr, ind_ii, L, U, face, o["v"], o["ncf"], o["P"].nf)
    e1, e0, e10 = rel(out[1], pv), rel(out[0], pv), rel(out[1], out[0])
    say("  rel err vs oracle: wave %.2e, 256-thread %.2e; between kernels %.2e" % (e1, e0, e10))
    assert e1 < 1e-5 and e0 < 1e-5 and e10 < 1e-5


def test_c2cfg_fp64_solve(c2cfg):
    from exasim_amd.hdg import HDGSolver
    say("case c2cfg: FP64 Newton solve")
    o = c2cfg
    with HDGSolver(o["path"], "convdiff") as S:
        S.solve()
        z = S.sizes
        uh = S.get("uh")
        udg = S.get("udg").reshape(z["ne"], z["nc"], z["npe"]).transpose(0, 2, 1)
    assert rel(uh, o["uh"]) < RTOL
    assert rel(udg, o["udg"]) < RTOL


def test_c2cfg_mixed_bf16_solve(c2cfg):
    """The benchmark's mixed leg: FP32 assembly and LU, BF16 block-ILU(0) factors, first solve
    and corrections to 1e-5, FP64 refinement to 1e-10 of the first residual."""
    from exasim_amd.hdg import HDGSolver
    say("case c2cfg: mixed form, BF16 block-ILU(0) factors")
    o = c2cfg
    with HDGSolver(o["path"], "convdiff") as S:
        S.set_option("precision", 1)
        S.set_option("precond_bf16", 1)
        S.set_option("refine_first_tol", 1e-5)
        S.set_option("refine_tol", 1e-5)
        it, gits, nrm = S.solve()
        z = S.sizes
        uh = S.get("uh")
        udg = S.get("udg").reshape(z["ne"], z["nc"], z["npe"]).transpose(0, 2, 1)
    err = rel(uh, o["uh"])
    say("  gmres %d, refinements %d, rel err vs FP64 oracle %.2e" % (gits, z["refinements"], err))
    assert err < 1e-8
    assert rel(udg, o["udg"]) < 1e-8


def _assembled(S, precond):
    S.set_option("preconditioner", precond)
    S.assemble()
    out = {k: S.get(k) for k in ("H", "DinvF", "Ru", "Rh", "b")}
    out["P"] = S.get({0: "BE", 1: "BEeas", 2: "BEilu"}[precond])
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_c2cfg_multiblock_assembly(c2cfg, precision):
    """A workspace of a few elements splits the 64 elements into blocks of neb (the last one
    partial): the element-block offsets of the residual, the boundary-face ranges, H, Rh and
    the mixed form's FP32 K', LU factors and pivots must give the one-block arrays."""
    from exasim_amd.hdg import HDGSolver
    o = c2cfg
    say("case c2cfg: multi-block assembly, precision %d" % precision)
    with HDGSolver(o["path"], "convdiff") as S:
        S.set_option("precision", precision)
        S.get_q()
        one = {p: _assembled(S, p) for p in (0, 1, 2)}
        assert S.state()["neb"] == S.sizes["ne"]
        per = 7.13e6  # bytes of one element's FP64 workspace at n = 250, m = 300 (FP32 form: 3.3e6)
        S.set_option("workspace_bytes", 9.5 * per)
        neb = S.state()["neb"]
        say("  %d elements per block" % neb)
        assert 1 < neb < S.sizes["ne"] and S.sizes["ne"] % neb != 0
        many = {p: _assembled(S, p) for p in (0, 1, 2)}
        nb, nru = S.residual()
        res_b = S.get("res_b")
        S.set_option("workspace_bytes", 16e9)
        S.residual()
        res_b1 = S.get("res_b")
    for p in (0, 1, 2):
        for k in one[p]:
            assert rel(many[p][k], one[p][k]) < 1e-13, (p, k)
    assert rel(res_b, res_b1) < 1e-13
    if precision == 0:
        assert rel(mats(many[2]["H"], 300, 300, 64), o["S"]["H"]) < RTOL
        assert rel(many[2]["b"], o["b"]) < RTOL


def test_c2cfg_multiblock_solve(c2cfg):
    from exasim_amd.hdg import HDGSolver
    o = c2cfg
    say("case c2cfg: FP64 Newton solve with a multi-block assembly")
    with HDGSolver(o["path"], "convdiff") as S:
        S.set_option("workspace_bytes", 9.5 * 7.13e6)
        assert S.state()["neb"] < S.sizes["ne"]
        S.solve()
        uh = S.get("uh")
    assert rel(uh, o["uh"]) < RTOL
