"""The served form (exasim_amd/serve.py): a timed stream of Navier-Stokes and
convection-diffusion requests of uneven size and order, admitted by HBM budget into
concurrent batches.  Every request completes, at least one batch holds several requests,
each result is bitwise the result of the same request solved alone, and each agrees with the FP64
oracle (oracle/hdg.py, the same datain) to 1e-8 relative, the served form's stated tolerance."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(tmp_path):
    from exasim_amd import cases, serve
    from exasim_amd.hdg import HDGSolver
    reqs = cases.serve_stream(str(tmp_path), seconds=0.4, rate=40.0, sizes=(4, 6), orders=(2, 3), ns_steps=1)
    print("served stream of %d requests" % len(reqs), flush=True)
    sys.stdout.flush()
    srv = serve.Server(hbm_budget_bytes=8e9, max_batch=4, max_concurrent=2, precision=1, keep_results=True,
                       window=0.2)
    summ = srv.run_stream(reqs)
    assert summ["completed"] == len(reqs) and summ["failed"] == 0, summ
    assert max(b[1] for b in srv.batches) > 1
    for rid, path, model, t in reqs:
        with HDGSolver.batch([path], model) as S:
            S.set_option("precision", srv.records[rid]["precision"])
            S.solve()
            uh = S.get_instance(0, "uh")
        assert np.array_equal(uh, srv.results[rid]), rid
    from oracle import hdg as ora
    worst = 0.0
    for rid, path, model, t in reqs:
        P = ora.HDGProblem(path, model)
        if P.tdep:
            ora.dirk_solve(P)
        else:
            ora.newton_solve(P)
        ref = P.uh.reshape(-1)
        err = np.abs(srv.results[rid] - ref).max() / np.abs(ref).max()
        worst = max(worst, err)
        assert err < 1e-8, (rid, model, err)
    print("served results vs FP64 oracle: largest relative error %.2e" % worst, flush=True)


def test_served_refills_equal_alone(tmp_path):
    """One batch slot at a time: every request of the slot's mesh that waits when the slot's
    request finishes is admitted into the running batch (exa_batch_refill), and each result is
    still bitwise the request solved alone."""
    from exasim_amd import cases, serve
    from exasim_amd.hdg import HDGSolver
    reqs = cases.serve_stream(str(tmp_path), seconds=0.4, rate=40.0, sizes=(4, 6), orders=(2, 3), ns_steps=1)
    print("served stream of %d requests, one slot" % len(reqs), flush=True)
    srv = serve.Server(hbm_budget_bytes=8e9, max_batch=1, max_concurrent=1, precision=1, keep_results=True)
    summ = srv.run_stream(reqs)
    print("refills: %d, batches: %d" % (summ["refills"], summ["batches"]), flush=True)
    assert summ["completed"] == len(reqs) and summ["failed"] == 0, summ
    assert summ["refills"] > 0
    assert any(r.get("refilled") for r in srv.records.values())
    for rid, path, model, t in reqs:
        with HDGSolver.batch([path], model) as S:
            S.set_option("precision", srv.records[rid]["precision"])
            S.solve()
            uh = S.get_instance(0, "uh")
        assert np.array_equal(uh, srv.results[rid]), rid
