import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (REPO, HERE, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import libtensor_amd as lt
    ctx = lt.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def host_ctx():
    import libtensor_amd as lt
    ctx = lt.Context(-1)
    yield ctx
    ctx.close()
