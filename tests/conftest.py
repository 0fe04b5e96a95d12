import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libbtrc_hip.so")


@pytest.fixture(scope="session")
def _session_ctx():
    import btrc_amd
    c = btrc_amd.Context(0)
    yield c
    c.close()


@pytest.fixture
def ctx(_session_ctx):
    """one context for the whole session; every test starts from the box
    filter, the default tuning and the cache switched off"""
    c = _session_ctx
    c.set_filter("box")
    c.set_tuning(closest_postpone_leaves=0, any_postpone_leaves=0, trace_waves_per_cu=20, launch_graph=1,
                 two_level_postpone=0, stage_blocks_per_cu=0, graph_bounces=0)
    c.nrc_configure(enabled=0, train_depth=0)
    yield c


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
