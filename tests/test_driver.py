"""The reference-structured host driver (gnss-1_amd/driver/gnss_main.cpp; main.cu:84-252's order:
allocate -> wavenumbers -> IC + forward + dealias -> stats(0) -> loop {step; stats every n; slices;
checkpoint} -> free) run as a program.

The driver calls only include/gnss.h, so the same source is linked twice: against the engine
(gnss-1_amd/bin/gnss_main) and against the oracle (oracle/gnss_main_oracle).  On the CPU the
oracle build is checked against the same calls made through the Python binding, and a restart from
its checkpoint files against the uninterrupted run; on the GPU the engine build's printed statistics
are checked against the oracle build's."""
import os
import subprocess

import pytest

from conftest import PKG, ROOT

ENGINE_DRIVER = os.path.join(PKG, "bin", "gnss_main")
ORACLE_DRIVER = os.path.join(ROOT, "oracle", "gnss_main_oracle")


def run_driver(exe, args, cwd=None):
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=600, cwd=cwd)
    assert r.returncode == 0, r.stderr
    rows = []
    for line in r.stdout.splitlines():
        if not line.startswith("step "):
            continue
        w = line.split()
        rows.append({"step": int(w[1]), "t": float(w[3]), "ke": float(w[5]), "epsilon": float(w[7]),
                     "chi0": float(w[9]), "cfl": float(w[11])})
    return rows, r.stdout


def close(a, b, tol):
    return abs(a - b) <= tol * abs(b)


def test_oracle_driver_follows_the_call_sequence(g, oracle):
    """Statistics before the first step and every 4 steps, identical to the same ABI calls made
    from Python (the driver adds nothing between them)."""
    rows, _ = run_driver(ORACLE_DRIVER, ["--n", "32", "--steps", "8", "--stats-every", "4", "--re", "800",
                                         "--sc", "0.7", "--theta", "0.08", "--kmax", "5", "--dt", "2e-3"])
    assert [r["step"] for r in rows] == [0, 4, 8]
    cfg = g.Config(n=(32,) * 3, re=800.0, sc=(0.7,), theta=0.08, noise_kmax=5.0, dt=2e-3, jet_h=1.0,
                   noise_amp=0.05, seed=20261202)
    s = g.Solver(cfg, oracle).setup()
    for k, r in enumerate(rows):
        if k:
            s.step(4)
        st = s.stats(False, False)[0]
        assert st["step"] == r["step"] and r["t"] == pytest.approx(st["t"], abs=1e-12)
        for q in ("ke", "epsilon", "cfl"):
            assert r[q] == st[q], (r["step"], q)
        assert r["chi0"] == st["chi"][0]


def test_oracle_driver_restarts_from_its_checkpoint(tmp_path):
    """--checkpoint-every writes the reference's raw field files; --restart continues from them and
    reaches the uninterrupted run's statistics (the RK3 integrator is self-starting, so fields and
    clock are the whole state; the physical-space round trip costs only roundoff)."""
    base = ["--n", "32", "--stats-every", "4", "--kmax", "5", "--theta", "0.08"]
    full, _ = run_driver(ORACLE_DRIVER, base + ["--steps", "8"])
    d = str(tmp_path)
    run_driver(ORACLE_DRIVER, base + ["--steps", "4", "--checkpoint-every", "4", "--out", d])
    assert os.path.exists(os.path.join(d, "u.4")) and os.path.exists(os.path.join(d, "z.4"))
    again, _ = run_driver(ORACLE_DRIVER, base + ["--steps", "4", "--restart", "4", "--out", d])
    assert [r["step"] for r in again] == [4, 8]
    for q in ("ke", "epsilon", "chi0", "cfl"):
        assert close(again[-1][q], full[-1][q], 1e-12), q
        assert close(again[0][q], full[1][q], 1e-12), q


@pytest.mark.gpu
def test_engine_driver_matches_oracle_driver():
    """The drop-in driver on the MI355X engine: the 64^3 jet (FP64), 20 steps, statistics every 10
    steps, against the same driver on the oracle (relative 1e-12)."""
    args = ["--config", "jet64", "--steps", "20", "--stats-every", "10"]
    print("\n[gnss_main --config jet64 --steps 20 --stats-every 10: engine and oracle builds]", flush=True)
    a, out = run_driver(ENGINE_DRIVER, args)
    b, _ = run_driver(ORACLE_DRIVER, args)
    print(out, flush=True)
    assert [r["step"] for r in a] == [r["step"] for r in b] == [0, 10, 20]
    for ra, rb in zip(a, b):
        for q in ("ke", "epsilon", "chi0", "cfl"):
            assert close(ra[q], rb[q], 1e-12), (ra["step"], q, ra[q], rb[q])
