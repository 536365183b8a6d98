"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

os.environ.setdefault("PDAL_QUIET", "1")

ORACLE_LP = ROOT / "oracle" / "_build" / "libcp_al_oracle.so"
PRODUCT_LP = ROOT / "pdal_amd" / "lib" / "libcp_al.so"
ORACLE_QP = ROOT / "oracle" / "_build" / "libfa_cp_oracle.so"
ORACLE_QP_REV = ROOT / "oracle" / "_build" / "libfa_cp_oracle_rev.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_lp():
    """CPU oracle library (test infrastructure only)."""
    from pdal_amd.lp import CpAl
    if not ORACLE_LP.exists():
        pytest.fail(f"oracle not built: run python oracle/build_oracle.py ({ORACLE_LP})")
    return CpAl(ORACLE_LP)


@pytest.fixture(scope="session")
def gpu_lp():
    """The MI355X product library; the test fails (no fallback) if it is missing."""
    from pdal_amd.lp import load_cp_al
    return load_cp_al()


@pytest.fixture(scope="session")
def oracle_qp():
    """CPU FA_CP oracle library (test infrastructure only)."""
    from pdal_amd.qp import FaCp
    if not ORACLE_QP.exists():
        pytest.fail(f"oracle not built: run python oracle/build_oracle.py ({ORACLE_QP})")
    return FaCp(ORACLE_QP)


@pytest.fixture(scope="session")
def oracle_qp_rev():
    """The FA_CP oracle built to sum in reverse order (same algorithm, other rounding)."""
    from pdal_amd.qp import FaCp
    if not ORACLE_QP_REV.exists():
        pytest.fail(f"oracle not built: run python oracle/build_oracle.py ({ORACLE_QP_REV})")
    return FaCp(ORACLE_QP_REV)


@pytest.fixture(scope="session")
def gpu_qp():
    """The MI355X FA_CP library; the test fails (no fallback) if it is missing."""
    from pdal_amd.qp import load_fa_cp
    return load_fa_cp()


@pytest.fixture(autouse=True)
def _clean_cp_al_env(monkeypatch):
    # tests set the reference's environment knobs explicitly
    for k in list(os.environ):
        if k.startswith("CP_AL_"):
            monkeypatch.delenv(k, raising=False)
    yield
