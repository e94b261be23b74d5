"""exa_matvec_noalloc -- the condensed matvec with its element products in the library's fixed
device array (no allocation) -- against exa_matvec on the same handle and inputs, bit for bit, in
FP64 and with the mixed form's FP32 operator.  Two calls in a row through the C ABI, the second on
a smaller problem; the first one split into element chunks (noalloc_chunk), so the chunked
accumulation is exercised, then the same handle again with whole-array chunks."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _matvecs(S, v):
    """(exa_matvec result, exa_matvec_noalloc result) for the same v."""
    px, _ = S.device_ptr("x")
    pw, _ = S.device_ptr("w")
    S.set("x", v)
    S.matvec(px, pw)
    ref = S.get("w")
    S.set("w", np.full_like(v, np.nan))
    rc = S._lib.exa_matvec_noalloc(S._h, ctypes.c_void_p(px), ctypes.c_void_p(pw))
    assert rc == 0, S._lib.exa_last_error(S._h).decode()
    return ref, S.get("w")


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matvec_equals_matvec(tmp_path, precision):
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver
    big = str(tmp_path / "big")
    small = str(tmp_path / "small")
    cases.convdiff3d(big, n=4, porder=3, ncu=2, restart=40, gmres_tol=1e-10)
    cases.poisson2d(small, n=3, porder=2, pgauss=4)
    rng = np.random.default_rng(23 + precision)
    for path, model, chunk in ((big, "convdiff", 7), (small, "poisson", 0)):
        say("exa_matvec_noalloc: %s, precision %d, chunk %d" % (model, precision, chunk))
        with HDGSolver(path, model) as S:
            S.set_option("precision", precision)
            S.set_option("noalloc_chunk", chunk)
            S.get_q()
            S.assemble()
            v = rng.standard_normal(S.sizes["N"])
            ref, got = _matvecs(S, v)
            assert np.isfinite(ref).all()
            assert np.array_equal(ref, got), np.abs(ref - got).max()
            if chunk:
                S.set_option("noalloc_chunk", 0)
                ref2, got2 = _matvecs(S, v)
                assert np.array_equal(ref2, ref) and np.array_equal(got2, ref)
