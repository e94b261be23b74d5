"""The datain writer against files the reference wrote (CPU only): the reference's
examples/Poisson/poisson3d datain (tests/golden/poisson3d.npz holds app.bin and mesh.bin verbatim)
is written again, byte for byte, by exasim_amd.datain from the values it holds: app.bin from the
problem's own settings over the writer's defaults (readpdeapp.cpp:360-440, 830-915), mesh.bin from
its element-to-vertex table, face permutation, boundary flags and conditions
(writebinaryfiles.cpp:155-263)."""
import os

import numpy as np

from exasim_amd import datain as dw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def blocks(a):
    n = int(a[0])
    sizes = a[1:1 + n].astype(int)
    off, out = 1 + n, []
    for k in sizes:
        out.append(a[off:off + k])
        off += k
    assert off == a.size
    return out


def test_app_bin_written_byte_for_byte(tmp_path):
    ref = np.load(os.path.join(ROOT, "tests", "golden", "poisson3d.npz"))["app"]
    path = str(tmp_path / "app.bin")
    # the poisson3d example's settings: preconditioner 1, GMRESiter 1000, physicsparam (1, 0), one
    # zero time step (steady); everything else is the writer's default
    dw.write_app(path, nd=3, ncu=1, flag={6: 1}, problem={14: 1000}, physicsparam=[1.0, 0.0], tau=[1.0],
                 solversparam=[1e-6, 1e-8, 1e-3, 0.0], dt=[0.0])
    got = np.fromfile(path, dtype=np.float64)
    for i, (g, r) in enumerate(zip(blocks(got), blocks(ref))):
        assert np.array_equal(g, r), (i, g, r)
    assert got.tobytes() == ref.astype(np.float64).tobytes()


def test_mesh_bin_written_byte_for_byte(tmp_path):
    ref = np.load(os.path.join(ROOT, "tests", "golden", "poisson3d.npz"))["mesh"]
    b = blocks(ref)
    ndims = b[0]
    nd, ne, nfe = int(ndims[0]), int(ndims[1]), int(ndims[4])
    perm = b[23].reshape(nfe, -1).astype(np.int64)          # perm[j + npf l]
    bf = b[24].reshape(ne, nfe)                               # bf[l + nfe e]
    t = b[26].reshape(ne, -1).astype(np.int64)                # t[v + nve e], 0-based
    bcs = b[27]
    assert np.array_equal(b[9], np.arange(ne)) and b[10][0] == ne     # one process
    path = str(tmp_path / "mesh.bin")
    dw.write_mesh(path, dict(nd=nd, nfe=nfe, perm=perm), t, bf, list(bcs))
    got = np.fromfile(path, dtype=np.float64)
    assert got.tobytes() == ref.astype(np.float64).tobytes()
