"""Shared test set-up: import paths, the `gpu` marker, built libraries."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "mpm-2d_amd")
ORACLE = os.path.join(REPO, "oracle")
for p in (PKG, ORACLE, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    """Builds the CPU oracle (and the HIP library) only if nothing is built."""
    if not os.path.exists(os.path.join(ORACLE, "_build", "liboracle.so")):
        subprocess.run(["make", "-C", ORACLE], check=True, capture_output=True)
    if not os.path.exists(os.path.join(PKG, "lib", "libmpm_hip.so")):
        subprocess.run(["make", "-C", PKG], check=True, capture_output=True)


_ensure_built()


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def pympm_mod():
    import pympm
    return pympm
