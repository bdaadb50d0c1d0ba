"""Shared fixtures: the engine library (GPU) and the FP64 oracle library (CPU, checker only)."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gnss-1_amd")
ORACLE_LIB = os.path.join(ROOT, "oracle", "libgnss_oracle.so")
ENGINE_LIB = os.path.join(PKG, "lib", "libgnss_hip.so")
HEADER = os.path.join(ROOT, "include", "gnss.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _binding():
    if "gnss" in sys.modules:
        return sys.modules["gnss"]
    spec = importlib.util.spec_from_file_location("gnss", os.path.join(PKG, "gnss.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gnss"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgnss_hip.so)")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def g():
    return _binding()


@pytest.fixture(scope="session")
def oracle(g):
    if not os.path.exists(ORACLE_LIB):
        pytest.fail("oracle/libgnss_oracle.so is not built (run __graft_entry__.build())")
    return g.load(ORACLE_LIB)


@pytest.fixture(scope="session")
def engine(g):
    # GPU tests must run the native engine: a missing library is a failure, never a skip.
    if not os.path.exists(ENGINE_LIB):
        pytest.fail("gnss-1_amd/lib/libgnss_hip.so is not built (run __graft_entry__.build())")
    return g.load(ENGINE_LIB)
