import os
import sys

import pytest

try:
    # PyTorch-ROCm bundles its own HIP runtime. When both are used in one
    # process, torch's must be loaded before libwhatprot_amd.so initialises
    # the system one, or torch finds no GPU afterwards.
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from tests import _oracle
    return _oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    from tests import _oracle
    r = _oracle.Reference.load()
    if r is None:
        pytest.skip("oracle/_ref/libwpref.so not built (needs the reference sources)")
    return r


@pytest.fixture(scope="session")
def gpu():
    import whatprot_amd as w
    if w.lib().wp_device_count() < 1:
        pytest.fail("no gfx950 device visible to the HIP runtime")
    return w
