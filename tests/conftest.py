import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU known-answer run (deselected by default with -m 'not slow')")


@pytest.fixture(scope="session")
def oracle():
    """ctypes handle of the oracle (CPU restatement, test infrastructure)."""
    from tests import _oracle
    return _oracle.load()


@pytest.fixture(scope="session")
def refcpu_cli():
    from tests import _oracle
    _oracle.load()
    return _oracle.CLI


@pytest.fixture(scope="session")
def gpu_ctx():
    import golaypair_amd as g
    ctx = g.Context(0)
    yield ctx
    ctx.close()
