"""Host logic of the served form (exasim_amd/serve.py) on the CPU: which requests count as solved,
which are solved again in FP64, and which are reported as failed.  The batch handles are replaced
by a stand-in whose final Newton residual norms are set per request and precision, so the test
exercises the server's bookkeeping only (the solves themselves are covered by the GPU tests)."""
import os
import shutil

import numpy as np
import pytest

from exasim_amd import cases, serve


class FakeBatch:
    """Stands in for HDGSolver.batch: every step finishes every instance, with the final
    residual norm NORMS[(request directory, precision)]."""
    NORMS = {}
    created = []

    def __init__(self, paths, model):
        self.paths = list(paths)
        self.prec = 1
        FakeBatch.created.append(self)

    @classmethod
    def batch(cls, paths, model):
        return cls(paths, model)

    def set_option(self, name, value):
        if name == "precision":
            self.prec = int(value)

    def step(self):
        return [0] * len(self.paths)

    def solve(self):
        return None

    def instance_stats(self):
        n = len(self.paths)
        norms = [FakeBatch.NORMS[(os.path.basename(p), self.prec)] for p in self.paths]
        return np.ones(n, dtype=np.int32), np.ones(n, dtype=np.int32), np.array(norms)

    def get_instance(self, i, name):
        return np.zeros(4)

    def refill(self, i, datain):
        self.paths[i] = datain

    def close(self):
        return None


@pytest.mark.parametrize("refill", [True, False])
def test_accept_fallback_and_failure(tmp_path, monkeypatch, refill):
    base = str(tmp_path / "base")
    cases.poisson2d(base, n=2, porder=1, pgauss=2)
    names = ["ok", "nan_then_ok", "nan_then_stuck", "loose_then_ok"]
    for nm in names:
        shutil.copytree(base, str(tmp_path / nm))
    FakeBatch.NORMS = {("ok", 1): 1e-12,
                       ("nan_then_ok", 1): float("nan"), ("nan_then_ok", 0): 1e-13,
                       ("nan_then_stuck", 1): float("nan"), ("nan_then_stuck", 0): 0.07,
                       ("loose_then_ok", 1): 1e-5, ("loose_then_ok", 0): 1e-12}
    FakeBatch.created = []
    monkeypatch.setattr(serve, "HDGSolver", FakeBatch)
    srv = serve.Server(hbm_budget_bytes=8e9, max_batch=4, max_concurrent=1, precision=1, refill=refill,
                       window=0.05)
    for k, nm in enumerate(names):
        srv.submit(k, str(tmp_path / nm), "poisson", arrival=0.0)
    srv.close()
    srv.join()
    s = srv.summary()
    assert s["completed"] == 3 and s["failed"] == 1, s
    assert s["fp64_fallback_requests"] == 3, s
    rec = srv.records
    assert rec[0]["precision"] == 1 and not rec[0].get("failed")
    assert rec[1]["precision"] == 0 and not rec[1].get("failed")
    assert rec[2].get("failed") and rec[2]["norm"] == pytest.approx(0.07)
    assert rec[3]["precision"] == 0 and rec[3]["norm"] == pytest.approx(1e-12)
    # the FP64 retry is a batch of the three requests the mixed form did not solve
    assert sorted(len(b.paths) for b in FakeBatch.created) == [3, 4]
