import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def ext():
    """The product module; GPU tests fail (not skip) when the library is missing."""
    from gplsaco_amd import gp_faco_ext
    gp_faco_ext.lib()
    return gp_faco_ext


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle


def announce(*parts):
    """Name the case on stdout before a GPU step (flushed)."""
    print("CASE", *parts, flush=True)
