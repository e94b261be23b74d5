"""GPU parity of model libraries: model functions in text2code's GPU form (the signatures of
backend/Model/libpdemodel.hpp, parallel_for bodies over points; exasim_amd/plugins/*) built with
include/CudaHip.hpp and loaded with the model name "plugin:<path>".

* The library for 3D convection-diffusion (two components) gives the built-in convdiff model's
  assembled system and solution, and the one for 2D Poisson the generated poisson2d model's, to
  1e-12 relative (the same formulas evaluated in another order).
* Both solutions agree with the oracle to 1e-10, including a multi-block assembly (a workspace of
  a few elements) and the mixed form (1e-8).
"""
import os
import sys

import numpy as np
import pytest

from oracle import hdg as ora

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUG = os.path.join(ROOT, "exasim_amd", "plugins")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


CASES = {
    "convdiff3d": ("convdiff3d", dict(n=2, porder=3, ncu=2, restart=100, gmres_tol=1e-12), "convdiff",
                   "libconvdiff3d_t2c.so"),
    "poisson2d": ("poisson2d", dict(n=4, porder=3, pgauss=8), "poisson2d", "libpoisson2d_t2c.so"),
}


def _system_and_solve(path, model, **opts):
    from exasim_amd.hdg import HDGSolver
    with HDGSolver(path, model) as S:
        for k, v in opts.items():
            S.set_option(k, v)
        S.get_q()
        S.assemble()
        sysm = {k: S.get(k) for k in ("H", "DinvF", "Ru", "Rh", "b")}
        S.reset()
        S.solve()
        return sysm, S.get("uh"), S.get("udg")


@pytest.mark.parametrize("name", list(CASES))
def test_model_library_matches_compiled_model(case_dir, name):
    builder, kw, twin, lib = CASES[name]
    so = os.path.join(PLUG, lib)
    assert os.path.exists(so), "model library not built: " + so
    say("case %s: model library %s against the model %s" % (name, lib, twin))
    path, info = case_dir("plugin_" + name, builder, **kw)
    ref_sys, ref_uh, ref_udg = _system_and_solve(path, twin)
    got_sys, got_uh, got_udg = _system_and_solve(path, "plugin:" + so)
    for k in ref_sys:
        assert rel(got_sys[k], ref_sys[k]) < 1e-12, k
    assert rel(got_uh, ref_uh) < 1e-12
    assert rel(got_udg, ref_udg) < 1e-12
    oracle_model = "poisson" if name == "poisson2d" else "convdiff"
    P = ora.HDGProblem(path, oracle_model)
    ora.newton_solve(P)
    assert rel(got_uh, P.uh.reshape(-1)) < 1e-10
    # several element blocks (the staging buffer sized per block) and the mixed form
    blk_sys, blk_uh, _ = _system_and_solve(path, "plugin:" + so, workspace_bytes=1e5)
    for k in ref_sys:
        assert rel(blk_sys[k], got_sys[k]) < 1e-12, k
    assert rel(blk_uh, P.uh.reshape(-1)) < 1e-10
    _, mix_uh, _ = _system_and_solve(path, "plugin:" + so, precision=1)
    assert rel(mix_uh, P.uh.reshape(-1)) < 1e-8
