"""The C-ABI libraries load and export every entry point include/gnss.h declares (no GPU calls)."""
import ctypes
import re

import pytest

from conftest import ENGINE_LIB, HEADER, ORACLE_LIB


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gnss_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_step_boundary():
    names = declared_functions()
    for must in ["gnss_create", "gnss_destroy", "gnss_init_wavenumbers", "gnss_init_jet", "gnss_step",
                 "gnss_stats", "gnss_forward", "gnss_inverse", "gnss_dealias", "gnss_get_field",
                 "gnss_set_field", "gnss_ensemble_create", "gnss_ensemble_step"]:
        assert must in names


def test_engine_exports_every_declared_symbol():
    lib = ctypes.CDLL(ENGINE_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_every_declared_symbol():
    lib = ctypes.CDLL(ORACLE_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header(g):
    assert set(declared_functions()) == set(g.SIGNATURES)


def test_oracle_rejects_bad_parameters(g, oracle):
    import pytest
    for bad in [dict(n=(10, 16, 16)), dict(dt=0.0), dict(re=-1.0), dict(sc=(0.7, 0.0))]:
        with pytest.raises(g.GnssError) as e:
            g.Solver(g.Config(**bad), oracle)
        assert e.value.code == 1


def test_oracle_reports_engine_only_calls_as_unsupported(g, oracle):
    s = g.Solver(g.Config(n=(16, 16, 16), sc=(0.7,)), oracle)
    s.init_wavenumbers()
    for call in (lambda: s.bench_fft(1), lambda: s.bench_copy(1), lambda: s.timing(True),
                 lambda: s.autotune(1)):
        with pytest.raises(g.GnssError) as e:
            call()
        assert e.value.code == 6  # GNSS_ERR_UNSUPPORTED
