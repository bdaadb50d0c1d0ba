"""Measured workloads against the oracle at reduced length.

jet256: plain FP64 drop-in at 256^3 with two scalars (Sc 0.7, 2.0), theta 0.03, Re 3000, noise
on |m| <= 16, dt 5e-4, statistics every 5 steps; 10 steps against the FP64 oracle (the full 200
steps run engine-only in gnss-1_amd/run_configs.py).  The oracle runs on the host's cores."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_jet256_two_scalars_vs_oracle(g, engine, oracle):
    print("\n[jet256 fp64 vs oracle, 10 steps]", flush=True)
    cfg = g.config("jet256")
    oracle.gnss_oracle_set_backend.argtypes = [g.ctypes.c_int]
    oracle.gnss_oracle_set_backend(1)  # FFTW3 API when present (same results, faster)
    try:
        a, b = g.Solver(cfg, engine).setup(), g.Solver(cfg, oracle).setup()
        for chunk in range(2):
            a.step(5)
            b.step(5)
            sa, ya, pa = a.stats()
            sb, yb, pb = b.stats()
            for k in ["ke", "epsilon", "enstrophy", "lint", "cfl", "umax"]:
                assert abs(sa[k] - sb[k]) <= 1e-11 * abs(sb[k]), k
            for s in range(2):
                assert abs(sa["chi"][s] - sb["chi"][s]) <= 1e-11 * sb["chi"][s]
            assert np.abs(ya - yb).max() <= 1e-11 * np.abs(yb).max()
            assert np.abs(pa - pb).max() <= 1e-11 * np.abs(pb).max()
            print(f"  step {sa['step']}: KE {sa['ke']:.10e} vs {sb['ke']:.10e}", flush=True)
        for f in range(5):
            assert rel(a.get_spectral(f), b.get_spectral(f)) < 1e-12, f
    finally:
        oracle.gnss_oracle_set_backend(0)
