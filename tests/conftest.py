"""Test configuration: the `gpu` marker and paths to the package and oracle."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tinker-hp_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import thp
    c = thp.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_stored():
    """A context whose FP32 systems store their matvec coefficients per solve
    (THP_MATVEC_RC=0 at thp_init) instead of recomputing them per entry."""
    import thp
    old = os.environ.get("THP_MATVEC_RC")
    os.environ["THP_MATVEC_RC"] = "0"
    try:
        c = thp.Context(0)
    finally:
        if old is None:
            del os.environ["THP_MATVEC_RC"]
        else:
            os.environ["THP_MATVEC_RC"] = old
    yield c
    c.close()
