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
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref built from the reference sources")


@pytest.fixture
def _ref_built():
    """Skips the test when oracle/_ref (the reference's own compilable sources) is not built.
    A fixture, so the library is loaded only by the tests that run, never at collection."""
    import _oracle as O
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")


needs_ref = pytest.mark.usefixtures("_ref_built")
