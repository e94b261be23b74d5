import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libexahip.so")


@pytest.fixture(scope="session")
def case_dir(tmp_path_factory):
    """Build (once per session) a datain directory for a named case."""
    from exasim_amd import cases
    made = {}

    def make(name, builder, **kw):
        if name not in made:
            path = str(tmp_path_factory.mktemp(name))
            info = getattr(cases, builder)(path, **kw)
            made[name] = (path, info)
        return made[name]
    return make


def make_poisson3d_datain(path):
    """Rebuild the reference's examples/Poisson/poisson3d datain from the golden
    fixture: app/master/mesh.bin verbatim; sol.bin holds xdg, the P2 nodes of the
    straight-sided tetrahedra (text2code makemesh.cpp:809-827 with the linear
    vertex basis of makemaster.cpp:848-938).  Returns the golden uhat sample."""
    import numpy as np
    from exasim_amd import datain as dw
    from oracle import hdg as ora
    z = np.load(os.path.join(ROOT, "tests", "golden", "poisson3d.npz"))
    os.makedirs(path, exist_ok=True)
    for k in ("app", "master", "mesh"):
        z[k].astype(np.float64).tofile(os.path.join(path, k + ".bin"))
    g = z["grid"]
    nd, npn, nve, ne = g[:4].astype(int)
    p = g[4:4 + nd * npn].reshape(npn, nd)
    t = (g[4 + nd * npn:].reshape(ne, nve) - 1).astype(np.int64)
    M = ora.read_master(os.path.join(path, "master.bin"))
    x = M["xpe"]
    phi = np.stack([1 - x[:, 0] - x[:, 1] - x[:, 2], x[:, 0], x[:, 1], x[:, 2]], 1)
    xdg = np.einsum("iv,evd->eid", phi, p[t])
    dw.write_sol(os.path.join(path, "sol.bin"), {"nfe": 4, "npf": M["npf"]}, xdg, ncu=1, nc=4)
    return z["faces"], z["uhat_faces"], z["uhat_header"]
