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
