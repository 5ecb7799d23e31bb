import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ncpaprop-release_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

REFERENCE_SAMPLES = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def synth_profile(tmp_path_factory):
    from ncpamodes import synth
    d = tmp_path_factory.mktemp("profiles")
    return synth.write_profile(str(d / "P0.dat"), seed=0)


@pytest.fixture(scope="session")
def gpu_ctx():
    import ncpamodes
    ctx = ncpamodes.Context(0)
    yield ctx
    ctx.close()
