import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (ROOT, os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture
def mutation():
    """Switch on one planted mistake in the oracle for the duration of a test."""
    import oracle

    def set_(k):
        oracle.lib().orc_set_mutation(int(k))

    yield set_
    oracle.lib().orc_set_mutation(0)
