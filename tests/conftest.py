"""Shared test setup: the `gpu` marker, import paths and library handles."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


def _ensure_oracle():
    so = os.path.join(ROOT, "oracle", "libstp_oracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)


_ensure_oracle()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """Session-wide HIP context (device 0) through the C-ABI."""
    from fastimaging_amd import stp
    ctx = stp.Context(0)
    yield ctx
    ctx.close()
