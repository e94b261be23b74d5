"""Regression: no kernel reads device memory before writing it.

Device memory freed by earlier work can hold any bit pattern; an update such as
y = a x + 0 y must not read y.  Each test first fills and frees a large device
allocation with 0xFF bytes (a NaN in FP64 and FP32), so the allocations the
library makes next are likely to reuse those pages, then runs a solve and checks
it against a solve made on a fresh handle.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def poison_device_memory(nbytes=2 << 30):
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(p, ctypes.c_int(0xFF), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipFree(p) == 0


@pytest.mark.parametrize("builder,kw,model,opts", [
    ("poisson2d", dict(n=4, porder=3, pgauss=8), "poisson2d", {}),
    ("convdiff_unsteady", dict(nd=2, n=4, porder=2, dt=0.05, nsteps=2, torder=3, nstage=3), "convdiff", {}),
    ("convdiff3d", dict(n=2, porder=2, ncu=2, gmres_tol=1e-12), "convdiff", {"precision": 1}),
])
def test_solve_after_poisoned_memory(case_dir, builder, kw, model, opts):
    print("case %s/%s %s after poisoning freed device memory" % (builder, model, opts), flush=True)
    sys.stdout.flush()
    from exasim_amd.hdg import HDGSolver
    path, _ = case_dir("poison_" + builder, builder, **kw)
    with HDGSolver(path, model) as S:
        for k, v in opts.items():
            S.set_option(k, v)
        S.solve()
        ref = S.get("uh")
    poison_device_memory()
    with HDGSolver(path, model) as S:
        for k, v in opts.items():
            S.set_option(k, v)
        S.solve()
        uh = S.get("uh")
    assert np.isfinite(uh).all()
    assert np.array_equal(uh, ref)
