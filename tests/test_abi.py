"""The C-ABI library loads and exports every entry point include/exasim_hdg.h declares
(no compute calls: these run without a GPU)."""
import os
import re

import pytest

from exasim_amd import hdg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "exasim_hdg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(exa_[a-z_]+|libmain)\s*\(", text)))


def test_header_matches_ctypes_table():
    assert declared_symbols() == sorted(n for n, _, _ in hdg.SIGNATURES)


def test_library_exports_every_symbol():
    if not os.path.exists(hdg.LIB_PATH):
        pytest.skip("libexahip.so not built (run python -m exasim_amd.build)")
    lib = hdg.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_create_reports_missing_input(tmp_path):
    """exa_create fails with EXA_ERR_IO before touching a device when app.bin is missing."""
    if not os.path.exists(hdg.LIB_PATH):
        pytest.skip("libexahip.so not built")
    lib = hdg.load()
    import ctypes
    h = ctypes.c_void_p()
    rc = lib.exa_create(str(tmp_path).encode(), None, b"poisson", -2, ctypes.byref(h))
    assert rc != 0
    assert h.value is None


DRIVER = os.path.join(ROOT, "exasim_amd", "gput2cEXASIM")


def test_driver_command_line_errors(tmp_path):
    """The gput2cEXASIM-style driver: usage, missing pdeapp.txt, missing model name and
    missing datain fail with a message and a non-zero status before any device work."""
    import subprocess
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built (run python -m exasim_amd.build)")
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([DRIVER, str(tmp_path / "none.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and "does not exist" in r.stderr
    app = tmp_path / "pdeapp.txt"
    app.write_text('model = "ModelD";\nmodelfile = "pdemodel.txt";\nporder = 3;\n')
    r = subprocess.run([DRIVER, str(app)], capture_output=True, text=True)
    assert r.returncode == 1 and "hipmodel" in r.stderr
    app.write_text('model = "ModelD";\nhipmodel = "poisson";\ndatapath = "%s";\n' % (tmp_path / "d"))
    r = subprocess.run([DRIVER, str(app)], capture_output=True, text=True)
    assert r.returncode == 2 and "Error" in r.stderr          # EXA_ERR_IO: no datain/app.bin
    assert os.path.isdir(tmp_path / "d" / "dataout")


def test_batch_create_checks_instances(tmp_path):
    """exa_batch_create rejects instances of different element shapes (EXA_ERR_ARG) and the
    block-ILU(0) preconditioner (EXA_ERR_UNSUPPORTED) before any device work."""
    if not os.path.exists(hdg.LIB_PATH):
        pytest.skip("libexahip.so not built")
    from exasim_amd import cases
    a, b, c = str(tmp_path / "a"), str(tmp_path / "b"), str(tmp_path / "c")
    cases.convdiff2d_param(a, n=2, porder=2)
    cases.convdiff2d_param(b, n=3, porder=3)
    cases.convdiff2d_param(c, n=2, porder=2, precond="bilu0")
    with pytest.raises(hdg.ExaError) as e:
        hdg.HDGSolver.batch([a, b], "cd2d", device=-2)
    assert e.value.code == 1 and "one master element" in str(e.value)
    with pytest.raises(hdg.ExaError) as e:
        hdg.HDGSolver.batch([c, c], "cd2d", device=-2)
    assert e.value.code == 5


def test_model_library_errors(tmp_path):
    """A model library ("plugin:<path>") that cannot be loaded, that lacks HdgFlux, or that is
    asked to serve a batch of several instances is refused before any device work."""
    if not os.path.exists(hdg.LIB_PATH):
        pytest.skip("libexahip.so not built")
    from exasim_amd import cases
    a = str(tmp_path / "a")
    cases.poisson2d(a, n=2, porder=2, pgauss=4)
    with pytest.raises(hdg.ExaError) as e:
        hdg.HDGSolver(a, "plugin:" + str(tmp_path / "missing.so"), device=-2)
    assert e.value.code == 1 and "cannot load model library" in str(e.value)
    with pytest.raises(hdg.ExaError) as e:
        hdg.HDGSolver(a, "plugin:" + hdg.LIB_PATH, device=-2)
    assert e.value.code == 1 and "has no HdgFlux" in str(e.value)
    lib = os.path.join(ROOT, "exasim_amd", "plugins", "libpoisson2d_t2c.so")
    if os.path.exists(lib):
        with pytest.raises(hdg.ExaError) as e:
            hdg.HDGSolver.batch([a, a], "plugin:" + lib, device=-2)
        assert e.value.code == 5
