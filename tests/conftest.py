import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ppf-registration-spatial-hashing_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libppf_amd.so)")


@pytest.fixture(scope="session")
def ctx():
    import ppf_amd
    c = ppf_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def batched_workload():
    """The batched configuration (16 models, 64 scenes), generated once."""
    import synth
    return synth.batched()
