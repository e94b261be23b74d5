"""Iteration-level batching on the GPU: exa_batch_step (one Newton iteration, or one DIRK time
step, of every instance still iterating) and exa_batch_refill (a finished slot takes a new
problem while the others go on).

* Steady nonlinear instances (wnl2d with different mu: several Newton iterations each, a w
  equation per point, one of them damped) stepped in one batch, finished slots refilled with new
  problems while the others iterate: each result is bitwise the result of the same problem alone
  (exa_solve on a batch of one) with the same Newton and GMRES counts, and agrees with the FP64
  oracle to 1e-10.
* The mixed form (cd2d requests of the served stream, FP32 factors with FP64 refinement): each
  result bitwise the mixed solve alone, and within 1e-8 of the FP64 oracle.
* DIRK(2,2) Navier-Stokes requests with one, two and three time steps: a slot refilled at a
  time-step boundary starts at the initial time while its neighbour is at a later one; each
  result is bitwise the problem solved alone.
* Refilling a slot that still iterates, or with another mesh, is refused.
"""
import os
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


def alone(path, model, precision=0):
    from exasim_amd.hdg import HDGSolver
    with HDGSolver.batch([path], model) as S:
        S.set_option("precision", precision)
        S.solve()
        stats = S.instance_stats()
        return S.get_instance(0, "uh"), S.get_instance(0, "udg"), (int(stats[0][0]), int(stats[1][0]))


def run_steps(first, pending, model, precision=0, max_steps=200):
    """Step a batch of the problems in ``first``; each finished slot takes the next of ``pending``.
    Returns {path: (uh, udg, (newton, gmres))} and the number of steps."""
    from exasim_amd.hdg import HDGSolver
    out = {}
    pending = list(pending)
    with HDGSolver.batch(list(first), model) as S:
        S.set_option("precision", precision)
        slots = list(first)
        for k in range(max_steps):
            st = S.step()
            newton, gmres, _ = S.instance_stats()
            for i, path in enumerate(slots):
                if path is None or st[i] == 1:
                    continue
                assert st[i] == 0, (path, st[i])
                out[path] = (S.get_instance(i, "uh"), S.get_instance(i, "udg"), (int(newton[i]), int(gmres[i])))
                slots[i] = None
                if pending:
                    slots[i] = pending.pop(0)
                    S.refill(i, slots[i])
            if all(p is None for p in slots):
                return out, k + 1
    raise AssertionError("the step loop did not finish")


def test_steady_steps_with_refill_equal_alone(tmp_path):
    from exasim_amd import cases
    mus = [(1.0, 1.0, 2.0), (0.5, 2.0, 1.0), (2.0, 0.5, 4.0), (1.0, 3.0, 0.5), (0.7, 1.0, 8.0)]
    paths = []
    for i, mu in enumerate(mus):
        p = str(tmp_path / ("w%d" % i))
        cases.wmodel2d(p, model="wnl2d", mu=mu, alpha=0.0, beta=1.0, u0=0.0, w0=0.0)
        paths.append(p)
    say("wnl2d steady: 3 slots, 2 refills")
    got, nsteps = run_steps(paths[:3], paths[3:], "wnl2d")
    say("  %d steps; Newton iterations %s" % (nsteps, [got[p][2][0] for p in paths]))
    assert nsteps > max(got[p][2][0] for p in paths)  # the refilled problems ran after the first ones
    for p in paths:
        uh, udg, counts = alone(p, "wnl2d")
        assert counts == got[p][2], (p, counts, got[p][2])
        assert np.array_equal(uh, got[p][0]) and np.array_equal(udg, got[p][1]), p
    for p in (paths[0], paths[3]):   # (the last one stops at the Newton iteration limit, damped)
        P = ora.HDGProblem(p, "wnl2d")
        ora.newton_solve(P)
        assert rel(got[p][0], P.uh.reshape(-1)) < 1e-10, p


def test_mixed_steps_with_refill_equal_alone(tmp_path):
    from exasim_amd import cases
    reqs = cases.serve_stream(str(tmp_path), seconds=0.3, rate=30.0, frac_ns=0.0, sizes=(6,), orders=(3,))
    paths = [r[1] for r in reqs][:5]
    assert len(paths) >= 4
    say("cd2d mixed: 2 slots, %d refills" % (len(paths) - 2))
    got, nsteps = run_steps(paths[:2], paths[2:], "cd2d", precision=1)
    for p in paths:
        uh, udg, counts = alone(p, "cd2d", precision=1)
        assert counts == got[p][2], (p, counts, got[p][2])
        assert np.array_equal(uh, got[p][0]), p
    P = ora.HDGProblem(paths[-1], "cd2d")
    ora.newton_solve(P)
    assert rel(got[paths[-1]][0], P.uh.reshape(-1)) < 1e-8


def test_dirk_steps_with_refill_equal_alone(tmp_path):
    from exasim_amd import cases
    paths = []
    for k, nst in enumerate((1, 3, 2)):
        reqs = cases.serve_stream(str(tmp_path / ("s%d" % k)), seconds=0.5, rate=40.0, frac_ns=1.0, sizes=(4,),
                                  orders=(2,), ns_steps=nst, seed=7 + k)
        paths.append(reqs[0][1])
    say("ns2d DIRK(2,2): slots of 1 and 3 time steps, the first refilled with a 2-step problem")
    got, nsteps = run_steps(paths[:2], paths[2:], "ns2d", precision=0)
    assert nsteps == 3
    for p in paths:
        uh, udg, counts = alone(p, "ns2d")
        assert counts == got[p][2], (p, counts, got[p][2])
        assert np.array_equal(uh, got[p][0]) and np.array_equal(udg, got[p][1]), p


def test_refill_refusals(tmp_path):
    from exasim_amd import cases
    from exasim_amd.hdg import ExaError, HDGSolver
    a, b, c = (str(tmp_path / x) for x in "abc")
    cases.wmodel2d(a, model="wnl2d", mu=(1.0, 1.0, 2.0), alpha=0.0, beta=1.0)
    cases.wmodel2d(b, model="wnl2d", mu=(0.5, 1.0, 2.0), alpha=0.0, beta=1.0)
    cases.wmodel2d(c, model="wnl2d", n=5, mu=(0.5, 1.0, 2.0), alpha=0.0, beta=1.0)
    with HDGSolver.batch([a], "wnl2d") as S:
        with pytest.raises(ExaError, match="has not finished"):
            S.refill(0, b)
        while S.step()[0] == 1:
            pass
        with pytest.raises(ExaError, match="mesh of the instance"):
            S.refill(0, c)
        S.refill(0, b)
        while S.step()[0] == 1:
            pass
        uh = S.get_instance(0, "uh")
    assert np.array_equal(uh, alone(b, "wnl2d")[0])
    assert os.path.isdir(c)
