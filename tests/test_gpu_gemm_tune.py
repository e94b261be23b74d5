"""The GEMM tile autotuner (gemm.hip): every workgroup tile runs the same MFMA instruction over the
same K order for each output element, so a solve gives bitwise the same result whichever tiles the
tuner picks; the tuner records each timed product shape with a valid tile and a time per tile."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", [0, 1])
def test_tiles_give_identical_solves(tmp_path, precision):
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver, gemm_tuning
    path = str(tmp_path / "cd3d")
    cases.convdiff3d(path, n=4, porder=3, restart=50, gmres_tol=1e-12, newton_tol=1e-10, ortho=1)
    out = {}
    for mode in (0, 1):
        say("convdiff3d 4^3 p=3, precision %d, gemm_autotune %d" % (precision, mode))
        with HDGSolver(path, "convdiff") as S:
            S.set_option("gemm_autotune", mode)
            S.set_option("precision", precision)
            S.solve()
            out[mode] = (S.get("uh"), S.get("H"), S.get("DinvF"))
            S.set_option("gemm_autotune", -1)
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    rows = gemm_tuning()
    assert rows, "no product shape was timed"
    tiles = {"64x64", "32x64", "64x32", "128x64", "64x128"}
    for r in rows:
        us = np.array(r["us"])          # -1: a tile the product cannot take (computed in place)
        assert r["tile"] in tiles and len(us) == 5 and us[0] > 0.0 and np.all((us > 0.0) | (us == -1.0)), r
        # the chosen tile is the fastest (times are reported to 1 ns, so two tiles may print alike)
        chosen = us[["64x64", "32x64", "64x32", "128x64", "64x128"].index(r["tile"])]
        assert chosen > 0.0 and chosen <= us[us > 0.0].min() + 1.5e-3, r
    say("tuned shapes: %d, tiles %s" % (len(rows), sorted({r["tile"] for r in rows})))


TILES = ["64x64", "32x64", "64x32", "128x64", "64x128"]


@pytest.mark.parametrize("porder", [2, 3, 4])
@pytest.mark.parametrize("precision", [0, 1])
def test_each_tile_bitwise(tmp_path, porder, precision):
    """Every product of the assembly, the condensation LU and the Schwarz blocks forced onto each
    tile in turn (a product computed in place keeps a tile that holds all its rows): the condensed
    operator, the preconditioner blocks and the solve are bitwise those of the default tile."""
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver
    path = str(tmp_path / "cd3d")
    cases.convdiff3d(path, n=3, porder=porder, ncu=2, precond="eas", restart=40, gmres_tol=1e-10,
                     newton_tol=1e-8, ortho=1)
    ref = None
    for t, name in enumerate(TILES):
        say("convdiff3d 3^3 p=%d precision %d, every GEMM on tile %s" % (porder, precision, name))
        with HDGSolver(path, "convdiff") as S:
            S.set_option("gemm_tile", t)
            try:
                S.set_option("precision", precision)
                S.get_q()
                S.assemble()
                arrs = [S.get(k) for k in ("H", "DinvF", "Rh", "BEeas")]
                S.solve()
                arrs.append(S.get("uh"))
            finally:
                S.set_option("gemm_tile", -1)
        if ref is None:
            ref = arrs
            continue
        for k, a, b in zip(("H", "DinvF", "Rh", "BEeas", "uh"), arrs, ref):
            assert np.array_equal(a, b), (name, k, float(np.abs(a - b).max()))


def test_each_tile_bitwise_mixed_batch(tmp_path):
    """A p = 4 batch of 3D convection-diffusion instances in the mixed form, whose FP32 Schwarz-block
    LU applies 64-row triangle inverses in place: every tile (a 32-row tile would split those
    in-place products) gives the default tile's solution bit for bit, with the same GMRES counts."""
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver
    inst = cases.convdiff3d_batch(str(tmp_path), count=3, orders=(4,), sizes=(6,))
    paths = [d["path"] for d in inst]
    ref = None
    for t, name in enumerate(TILES):
        say("convdiff3d batch of 3 (6^3, p=4), mixed, every GEMM on tile %s" % name)
        with HDGSolver.batch(paths, "convdiff") as S:
            S.set_option("gemm_tile", t)
            try:
                S.set_option("precision", 1)
                S.set_option("refine_first_tol", 1e-5)
                S.set_option("refine_tol", 1e-5)
                S.solve()
                out = ([S.get_instance(i, "uh") for i in range(len(paths))], list(S.instance_stats()[1]))
            finally:
                S.set_option("gemm_tile", -1)
        if ref is None:
            ref = out
            continue
        assert out[1] == ref[1], (name, out[1], ref[1])
        for a, b in zip(out[0], ref[0]):
            assert np.array_equal(a, b), name
