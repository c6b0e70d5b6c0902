"""Shared fixtures: paths, the oracle front end and the GPU package.

Tests marked `gpu` call the HIP library through its C-ABI and need an MI355X;
everything else runs on the CPU (oracle against the reference's fixtures,
host logic, and that the C-ABI library loads and exports its symbols).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "fpga-litho-accel_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    import litho_amd

    litho_amd.lib()
    return litho_amd


def golden(name):
    return os.path.join(GOLDEN, name)


def max_rel(a, b):
    """max|a-b| / max|b|, the reference harness's image metric
    (validation/golden/run_verification.py L203-204)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a.astype(np.complex128) - b.astype(np.complex128)).max() /
                 max(np.abs(b).max(), 1e-300))


# Production-clip optics: lambda 193 nm, NA 0.85, zero defocus, 41x41 annulus
# sigma 0.6-0.9, 4 nm pixels (period = 4 N).
BASE_LAMBDA = 193.0
BASE_NA = 0.85


@pytest.fixture(scope="session")
def base_source(orc):
    src = orc.source_annular(41, 0.6, 0.9)
    return src, orc.outer_sigma(src)


# The reference's own fixture optics (source/TCC_HLS golden, kernel_info
# resolution sweep): 101x101 annulus 0.6-0.9, NA 0.8, defocus 0.2, 1 nm pixels.
FIX_NA = 0.8
FIX_DEFOCUS = 0.2


@pytest.fixture(scope="session")
def fix_source(orc):
    src = orc.source_annular(101, 0.6, 0.9)
    return src, orc.outer_sigma(src)
