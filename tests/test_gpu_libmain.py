"""The reference library's entry point, libmain(argc, argv) (backend/Main/libexasim.h:5), called
through the C ABI with the reference CI's arguments "Poisson3D 1 datain/ dataout/"
(.github/workflows/exasim-ci.yml:49-54) on the reference's examples/Poisson/poisson3d inputs: the
written dataout/uhat_np0.bin reproduces the reference's committed outuhat_np0.bin (both solved with
the GMRES tolerance of app.bin, 1e-8) to 5e-8, and bad arguments return error codes."""
import ctypes
import os

import numpy as np
import pytest

from conftest import make_poisson3d_datain

pytestmark = pytest.mark.gpu


def _argv(args):
    arr = (ctypes.c_char_p * len(args))()
    arr[:] = [a.encode() for a in args]
    return len(args), arr


def test_libmain_poisson3d(tmp_path):
    from exasim_amd import hdg
    from exasim_amd.datain import read_output
    from test_golden_poisson3d import check
    lib = hdg.load()
    datain = tmp_path / "datain"
    faces, ref, hdr = make_poisson3d_datain(str(datain))
    os.makedirs(tmp_path / "dataout")
    print("libmain Poisson3D 1 datain/ dataout/", flush=True)
    argc, argv = _argv(["libmain", "Poisson3D", "1", str(datain) + "/", str(tmp_path / "dataout") + "/"])
    assert lib.libmain(argc, argv) == 0
    h, uh = read_output(str(tmp_path / "dataout" / "uhat_np0.bin"))
    assert [int(v) for v in h] == [int(v) for v in hdr]
    uh = uh.reshape(int(h[2]), -1)
    h2, udg = read_output(str(tmp_path / "dataout" / "udg_np0.bin"))
    assert int(h2[2]) == 24576
    # interior faces come first in the face numbering; the boundary ones hold Dirichlet zeros
    from oracle import hdg as ora
    P = ora.HDGProblem(str(datain), "poisson")
    check(uh, P.conn["nif"], faces, ref)


def test_libmain_rejects_bad_arguments(tmp_path):
    from exasim_amd import hdg
    lib = hdg.load()
    assert lib.libmain(*_argv(["libmain", "Poisson3D", "1"])) == 1          # too few arguments
    assert lib.libmain(*_argv(["libmain", "Poisson3D", "2", "a/", "b/", "c/", "d/"])) == 5  # two models
    assert lib.libmain(*_argv(["libmain", "Poisson3D", "1", "a/", "b/", "3"])) == 5         # a restart step
    assert lib.libmain(*_argv(["libmain", "Poisson3D", "x", "a/", "b/"])) == 1             # bad nummodels
    assert lib.libmain(*_argv(["libmain", "Poisson3D", "1", str(tmp_path) + "/", "x/"])) == 2  # no app.bin
    assert lib.libmain(*_argv(["libmain", "NoSuchModel", "1", str(tmp_path) + "/", "x/"])) == 1
