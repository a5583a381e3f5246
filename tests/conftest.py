import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu_renderer_factory():
    """Creates renderers on cuda:0 and exits them at session end."""
    from wavefrontpathtracing_amd import Renderer

    made = []

    def make(w=64, h=64):
        r = Renderer(w, h)
        made.append(r)
        return r

    yield make
    for r in made:
        r.exit()
