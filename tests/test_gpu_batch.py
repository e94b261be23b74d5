"""Batches of independent instances (exa_batch_create) on the GPU.

* Each instance's result in a batch is bitwise the result of the same instance alone (a batch
  of one): the shared launches are element- and face-local and every Krylov/Newton scalar is
  the instance's own (segmented reductions over its own unknowns).
* Against the single-instance path (exa_create) and the oracle: 1e-10 relative (GMRES 1e-12).
* Mixed precision batch against the FP64 batch: 1e-8 relative.
* A nonlinear batch (two Navier-Stokes instances with different Mach and Reynolds numbers,
  DIRK(2,2) steps) equals each instance alone, bitwise.
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


INSTANCES = [dict(n=4, kappa=0.3, angle=0.4, k=1), dict(n=6, kappa=0.01, angle=2.0, k=3),
             dict(n=8, kappa=0.002, angle=4.1, k=2), dict(n=5, kappa=1.0, angle=5.5, k=4)]


@pytest.fixture(scope="module")
def cd_batch(tmp_path_factory):
    from exasim_amd import cases
    root = tmp_path_factory.mktemp("cdbatch")
    paths = []
    for i, kw in enumerate(INSTANCES):
        p = str(root / ("i%d" % i))
        cases.convdiff2d_param(p, porder=3, restart=40, gmres_tol=1e-12, newton_tol=1e-10, **kw)
        paths.append(p)
    return paths


def _solve_batch(paths, model="cd2d", precision=0):
    from exasim_amd.hdg import HDGSolver
    with HDGSolver.batch(paths, model) as S:
        assert S.ninst == len(paths)
        S.set_option("precision", precision)
        S.solve()
        stats = S.instance_stats()
        out = [(S.get_instance(i, "uh"), S.get_instance(i, "udg")) for i in range(len(paths))]
    return out, stats


def test_batch_equals_alone_bitwise(cd_batch):
    say("cd2d batch of 4 (p=3, n = 4..8) vs each instance alone")
    together, stats = _solve_batch(cd_batch)
    assert (stats[0] == 1).all() and (stats[1] > 0).all() and (stats[2] < 1e-10).all()
    for i, p in enumerate(cd_batch):
        alone, st1 = _solve_batch([p])
        assert np.array_equal(together[i][0], alone[0][0]), i
        assert np.array_equal(together[i][1], alone[0][1]), i
        assert st1[1][0] == stats[1][i]


def test_batch_matches_single_path_and_oracle(cd_batch):
    say("cd2d batch vs exa_create path and vs oracle")
    from exasim_amd.hdg import HDGSolver
    together, _ = _solve_batch(cd_batch)
    for i, p in enumerate(cd_batch):
        with HDGSolver(p, "cd2d") as S:
            S.solve()
            uh = S.get("uh")
        assert rel(together[i][0], uh) < 1e-10, i
    P = ora.HDGProblem(cd_batch[0], "cd2d")
    ora.newton_solve(P)
    assert rel(together[0][0], P.uh.reshape(-1)) < 1e-10


def test_batch_mixed_precision(cd_batch):
    say("cd2d batch: mixed precision vs FP64")
    f64, _ = _solve_batch(cd_batch)
    mix, stats = _solve_batch(cd_batch, precision=1)
    for i in range(len(cd_batch)):
        assert rel(mix[i][0], f64[i][0]) < 1e-8, i
        assert rel(mix[i][1], f64[i][1]) < 1e-8, i


def test_ns_batch_equals_alone(tmp_path):
    say("ns2d batch of 2 (M = 0.5 / Re 1000, M = 0.3 / Re 400), DIRK(2,2) x 2 vs alone")
    from exasim_amd import cases
    paths = []
    for i, (mach, re) in enumerate([(0.5, 1000.0), (0.3, 400.0)]):
        p = str(tmp_path / ("ns%d" % i))
        cases.navierstokes2d(p, nx=8 + 4 * i, ny=2, porder=3, length=(2.5, 1.0), mach=mach, re=re, nsteps=2,
                             torder=2, nstage=2, precond="bjacobi", restart=100, gmres_tol=1e-12,
                             newton_tol=1e-10, cartgridpart=False)
        paths.append(p)
    together, stats = _solve_batch(paths, "ns2d")
    assert (stats[0] >= 4).all()
    for i, p in enumerate(paths):
        alone, _ = _solve_batch([p], "ns2d")
        assert np.array_equal(together[i][0], alone[0][0]), i
        assert np.array_equal(together[i][1], alone[0][1]), i


def test_batch_isolates_a_failed_instance(cd_batch, tmp_path):
    say("cd2d batch with one instance whose diffusivity is NaN: the others are unaffected")
    from exasim_amd import cases
    bad = str(tmp_path / "bad")
    cases.convdiff2d_param(bad, n=4, porder=3, kappa=float("nan"), restart=40, gmres_tol=1e-12,
                           newton_tol=1e-10)
    out, stats = _solve_batch([cd_batch[0], bad, cd_batch[1]])
    norms = stats[2]
    assert np.isnan(norms[1]) and np.isfinite(norms[0]) and np.isfinite(norms[2])
    for k, p in ((0, cd_batch[0]), (2, cd_batch[1])):
        alone, _ = _solve_batch([p])
        assert np.array_equal(out[k][0], alone[0][0]), k


@pytest.fixture(scope="module")
def cd_batch_cgs(tmp_path_factory):
    """The batch workload's solver settings (elemental Schwarz, classical Gram-Schmidt,
    GMRES tolerance 1e-10, Newton 1e-8) with a restart of 5, so that restarts and the
    segmented multi-vector dot of four and more vectors are exercised."""
    from exasim_amd import cases
    root = tmp_path_factory.mktemp("cdbatch_cgs")
    paths = []
    for i, kw in enumerate(INSTANCES):
        p = str(root / ("i%d" % i))
        cases.convdiff2d_param(p, porder=3, precond="eas", restart=5, gmres_tol=1e-10, newton_tol=1e-8,
                               ortho=1, **kw)
        paths.append(p)
    return paths


def test_batch_cgs_matches_oracle(cd_batch_cgs):
    say("cd2d batch, CGS, Schwarz, GMRES(5): every instance vs the oracle (FP64 1e-10, mixed 1e-8)")
    f64, stats = _solve_batch(cd_batch_cgs)
    mix, _ = _solve_batch(cd_batch_cgs, precision=1)
    for i, p in enumerate(cd_batch_cgs):
        P = ora.HDGProblem(p, "cd2d")
        assert P.ortho == 1 and P.restart == 5
        hist = ora.newton_solve(P)
        ref = P.uh.reshape(-1)
        say("  instance %d: gmres gpu %d oracle %d" % (i, stats[1][i], sum(hist["gmres_iters"])))
        assert stats[1][i] == sum(hist["gmres_iters"]), i
        assert rel(f64[i][0], ref) < 1e-10, i
        udg = f64[i][1].reshape(P.ne, P.nc, P.npe).transpose(0, 2, 1)
        assert rel(udg, P.udg) < 1e-10, i
        assert rel(mix[i][0], ref) < 1e-8, i


def test_batch_mixed_refine_cycles_matches_oracle(cd_batch_cgs):
    """Mixed form with every GMRES solve capped at one restart cycle (refine_cycles = 1, GMRES(5)):
    the refinement corrections carry the convergence, per instance, to the FP64 oracle's solution
    within 1e-8."""
    say("cd2d batch, mixed, refine_cycles 1 with GMRES(5): every instance vs the oracle (1e-8)")
    from exasim_amd.hdg import HDGSolver
    with HDGSolver.batch(cd_batch_cgs, "cd2d") as S:
        for k, v in dict(precision=1, refine_cycles=1, refine_maxit=200, refine_first_tol=1e-7,
                         refine_tol=1e-7).items():
            S.set_option(k, v)
        S.solve()
        a, g, nrm = S.instance_stats()
        mix = [S.get_instance(i, "uh") for i in range(len(cd_batch_cgs))]
    for i, p in enumerate(cd_batch_cgs):
        P = ora.HDGProblem(p, "cd2d")
        ora.newton_solve(P)
        say("  instance %d: gmres %d, residual %.2e" % (i, g[i], nrm[i]))
        assert rel(mix[i], P.uh.reshape(-1)) < 1e-8, i


def test_batch3d_convdiff_matches_oracle(tmp_path):
    """A batch of 3D convection-diffusion instances (the 3D batch workload's settings: Schwarz blocks,
    classical Gram-Schmidt, two components, per-instance velocity, diffusivity and source, Dirichlet
    inflow faces) against
    the oracle: FP64 1e-10, mixed form with the benchmark's refinement tolerances 1e-8."""
    say("convdiff3d batch of 3 (n = 2, 3, 2; p = 2) vs oracle")
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver
    paths = []
    for i, (n, kappa, vel) in enumerate([(2, 0.05, (1.0, 0.0, 0.0)), (3, 0.01, (0.0, 0.6, 0.8)),
                                         (2, 0.1, (0.48, 0.6, 0.64))]):
        p = str(tmp_path / ("i%d" % i))
        cases.convdiff3d(p, n=n, porder=2, ncu=2, kappa=kappa, vel=vel, seed=7 + i, precond="eas",
                         restart=100, gmres_tol=1e-10, newton_tol=1e-8, ortho=1, inflow_dirichlet=True)
        paths.append(p)
    f64, stats = _solve_batch(paths, "convdiff")
    with HDGSolver.batch(paths, "convdiff") as S:
        S.set_option("precision", 1)
        S.set_option("refine_first_tol", 1e-5)
        S.set_option("refine_tol", 1e-5)
        S.solve()
        mix = [S.get_instance(i, "uh") for i in range(len(paths))]
    for i, p in enumerate(paths):
        P = ora.HDGProblem(p, "convdiff")
        hist = ora.newton_solve(P)
        ref = P.uh.reshape(-1)
        assert max(hist["gmres_iters"]) < 2000, "the oracle did not converge (instance %d)" % i
        assert stats[1][i] == sum(hist["gmres_iters"]), i
        assert rel(f64[i][0], ref) < 1e-10, i
        assert rel(mix[i], ref) < 1e-8, i
