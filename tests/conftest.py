import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libm3dref.so")
REF_LIBM_LIB = os.path.join(ROOT, "oracle", "_ref", "libm3dref_libm.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "c", "liboracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")


@pytest.fixture(scope="session")
def ref():
    """The reference engine built from its own sources (strict IEEE)."""
    from melon3d_amd.api import load_library

    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (reference tree absent when build() ran)")
    return load_library(REF_LIB)


@pytest.fixture(scope="session")
def ref_libm():
    """The reference engine built from its own sources on the system libm
    (its sinf, cosf, sincosf and atan2f untouched)."""
    from melon3d_amd.api import load_library

    if not os.path.exists(REF_LIBM_LIB):
        pytest.skip("oracle/_ref not built (reference tree absent when build() ran)")
    return load_library(REF_LIBM_LIB)


@pytest.fixture(scope="session")
def dev():
    """The MI355X backend; a missing build is an error, not a skip."""
    from melon3d_amd.api import load

    return load()
