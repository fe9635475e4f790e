import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def ctx():
    import accelgraph_express_amd as ag
    c = ag.Context(0)
    yield c
    c.close()


class tuned:
    """Context manager: library tuning knobs (ag.set_tuning) for gather
    layouts built inside the block, restored afterwards."""

    def __init__(self, **kw):
        self.kw = kw

    def __enter__(self):
        import accelgraph_express_amd as ag
        self.saved = ag.get_tuning()
        ag.set_tuning(**self.kw)
        return self

    def __exit__(self, *exc):
        import accelgraph_express_amd as ag
        ag.set_tuning(**self.saved)
