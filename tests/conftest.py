"""Shared test setup: marker registration, import paths and the checker build."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device and libmmprop_hip.so")
    config.addinivalue_line("markers", "slow: long-running (seconds to minutes)")
    oracle = os.path.join(REPO, "oracle", "_build", "libmmprop_oracle.so")
    if not os.path.exists(oracle):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)


def pytest_report_header(config):
    return f"mmprop tests, repo {REPO}"


@pytest.fixture(scope="session")
def capi_lib():
    from mm_propagation_amd import _capi
    return _capi.load_library()
