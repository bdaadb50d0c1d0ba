"""The five measured workloads at their stated shapes, on the GPU.

  jet64     plain FP64, 64^3, 100 steps: engine against the oracle, both timed over the step loop
  jet256    plain FP64, 256^3, two scalars (Sc 0.7, 2.0), 200 steps, statistics, planar profiles
            and the shell spectrum every 20 steps, engine against the oracle on the same seed
  jet512    the FP32 precision form at 512^3, 100 steps, against the FP64 plain form on the same
            seed: velocity/scalar rel-L2 <= 1e-5, integral statistics <= 1e-4, divergence <= 1e-6
  ensemble  32 x 64^3 + 16 x 96^3 + 8 x 128^3 FP32 members in one hipGraph per step, 200 steps:
            every member against its own FP64 run (<= 1e-5 after 100 and 200 steps), one member per grid
            through its FP64 run against the oracle
  jet1024   FP32 at 1024^3 with the split-FP16 MFMA radix-16 stage against the VALU stages, and
            both against FP64 at 512^3 with the same physical parameters on the 512^3 modes

The oracle (threaded FFTW3 API where present) runs on the host's cores as the checker.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _fftw(oracle, g, on):
    oracle.gnss_oracle_set_backend.argtypes = [g.ctypes.c_int]
    oracle.gnss_oracle_set_backend(1 if on else 0)


def _oracle_desc(g, oracle):
    oracle.gnss_oracle_backend_name.restype = g.ctypes.c_char_p
    oracle.gnss_oracle_max_threads.restype = g.ctypes.c_int
    return f"FFT {oracle.gnss_oracle_backend_name().decode()}, {oracle.gnss_oracle_max_threads()} threads"


@pytest.mark.timeout(300)
def test_jet64_full_configuration_timed(g, engine, oracle):
    """The whole 64^3 configuration (100 FP64 steps), engine and oracle timed over the step loop
    only, fields compared (rel-L2 <= 1e-12)."""
    print("\n[jet64 fp64, 100 steps, engine and oracle timed]", flush=True)
    cfg = g.config("jet64")
    _fftw(oracle, g, True)
    try:
        a, b = g.Solver(cfg, engine).setup(), g.Solver(cfg, oracle).setup()
        t0 = time.perf_counter()
        a.step(100)
        t1 = time.perf_counter()
        b.step(100)
        t2 = time.perf_counter()
        fr = [rel(a.get_spectral(f), b.get_spectral(f)) for f in range(4)]
        print(f"  engine {100 / (t1 - t0):.1f} steps/s, oracle {100 / (t2 - t1):.2f} steps/s "
              f"({_oracle_desc(g, oracle)}); fields rel-L2 {', '.join('%.1e' % x for x in fr)}", flush=True)
        assert max(fr) <= 1e-12, fr
    finally:
        _fftw(oracle, g, False)


# The 256^3 configuration is split into tests of 40 steps that continue one pair of runs (engine and
# oracle), so that each test ends, and reports, within a minute of host time.
@pytest.fixture(scope="module")
def jet256_runs(g, engine, oracle):
    print("\n[jet256 fp64 vs oracle, 200 steps, statistics every 20]", flush=True)
    cfg = g.config("jet256")
    _fftw(oracle, g, True)
    runs = {"a": g.Solver(cfg, engine).setup(), "b": g.Solver(cfg, oracle).setup(), "ta": 0.0, "tb": 0.0,
            "worst": {"stats": 0.0, "profiles": 0.0, "spectrum": 0.0}, "step": 0}
    yield runs
    _fftw(oracle, g, False)


def _jet256_compare(runs):
    a, b = runs["a"], runs["b"]
    sa, ya, pa = a.stats()
    sb, yb, pb = b.stats()
    ds = max(abs(sa[k] - sb[k]) / abs(sb[k]) for k in ["ke", "vrms", "epsilon", "enstrophy", "lint", "cfl", "umax"])
    ds = max([ds] + [abs(sa["chi"][s] - sb["chi"][s]) / sb["chi"][s] for s in range(2)])
    dy = float(np.abs(ya - yb).max() / np.abs(yb).max())
    dp = float(np.abs(pa - pb).max() / np.abs(pb).max())
    w = runs["worst"]
    runs["worst"] = {"stats": max(w["stats"], ds), "profiles": max(w["profiles"], dy), "spectrum": max(w["spectrum"], dp)}
    print(f"  step {sa['step']:3d}: KE {sa['ke']:.12e}  stats {ds:.1e}  profiles {dy:.1e}  spectrum {dp:.1e}", flush=True)
    assert sa["step"] == sb["step"] == runs["step"]
    assert ds <= 1e-11 and dy <= 1e-11 and dp <= 1e-11, (runs["step"], ds, dy, dp)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("chunk", range(6))
def test_jet256_full_configuration_vs_oracle(jet256_runs, chunk):
    """The whole 256^3 configuration: 200 FP64 steps with statistics, planar profiles and the
    shell spectrum every 20 steps, engine against the oracle on the same seed (statistics 1e-11
    relative; after 200 steps the fields to rel-L2 1e-12).  Chunk 0 compares the initial state,
    chunks 1-5 advance 40 steps each."""
    runs = jet256_runs
    if chunk == 0:
        _jet256_compare(runs)
        return
    for _ in range(2):
        t0 = time.perf_counter()
        runs["a"].step(20)
        t1 = time.perf_counter()
        runs["b"].step(20)
        runs["ta"] += t1 - t0
        runs["tb"] += time.perf_counter() - t1
        runs["step"] += 20
        _jet256_compare(runs)
    if chunk == 5:
        a, b, w = runs["a"], runs["b"], runs["worst"]
        fr = [rel(a.get_spectral(f), b.get_spectral(f)) for f in range(5)]
        print(f"  fields after 200 steps, rel-L2 (u, v, w, phi0, phi1): {', '.join('%.1e' % x for x in fr)}; "
              f"worst stats {w['stats']:.1e}, profiles {w['profiles']:.1e}, spectrum {w['spectrum']:.1e}", flush=True)
        print(f"  200 steps (statistics excluded): engine {200 / runs['ta']:.1f} steps/s, oracle "
              f"{200 / runs['tb']:.3f} steps/s", flush=True)
        assert max(fr) <= 1e-12, fr


@pytest.mark.timeout(600)
def test_jet512_fp32_form_vs_fp64(g, engine):
    """The precision form at its stated shape: 512^3, Re 5000, theta 0.02, Sc 1.0, noise 0.05 on
    |m| <= 32, dt 2.5e-4, 100 steps, FP32 transforms/products/updates with FP64 reductions, against
    the FP64 plain form on the same seed.  Bounds: velocity and scalar rel-L2 <= 1e-5, integral
    statistics <= 1e-4, FP32 post-projection divergence <= 1e-6; profiles and spectrum are held to
    the statistics bound (rows that vanish in exact arithmetic, such as <v>, by an absolute 1e-8)."""
    print("\n[jet512 fp32 form vs fp64 plain form, 100 steps]", flush=True)
    res = {}
    for name, prec in (("fp64", g.FP64), ("fp32", g.FP32)):
        s = g.Solver(g.config("jet512", precision=prec), engine).setup()
        st = []
        for k in range(2):
            s.step(50)
            st.append(s.stats())
        res[name] = (st, [s.get_spectral(f) for f in range(4)])
        print(f"  {name}: KE {st[-1][0]['ke']:.10e}, divergence {st[-1][0]['divergence']:.2e}", flush=True)
        s.close()
    keys = ("ke", "vrms", "epsilon", "enstrophy", "cfl", "umax", "lint", "lambda", "eta")
    for k in range(2):
        (a, ya, pa), (b, yb, pb) = res["fp32"][0][k], res["fp64"][0][k]
        ds = {q: abs(a[q] - b[q]) / abs(b[q]) for q in keys}
        ds["chi0"] = abs(a["chi"][0] - b["chi"][0]) / abs(b["chi"][0])
        ds["var0"] = abs(a["scalar_var"][0] - b["scalar_var"][0]) / abs(b["scalar_var"][0])
        big = [r for r in range(yb.shape[0]) if np.abs(yb[r]).max() > 1e-8]
        dy = max(rel(ya[r], yb[r]) for r in big)
        dz = max((float(np.abs(ya[r]).max()) for r in range(yb.shape[0]) if r not in big), default=0.0)
        dp = rel(pa, pb)
        print(f"  step {a['step']}: worst statistic {max(ds.values()):.1e} ({max(ds, key=ds.get)}), profiles "
              f"{dy:.1e} (zero rows {dz:.1e}), spectrum {dp:.1e}, fp32 divergence {a['divergence']:.1e}", flush=True)
        assert max(ds.values()) <= 1e-4, ds
        assert dy <= 1e-4 and dz <= 1e-8 and dp <= 1e-4
        assert a["divergence"] <= 1e-6
    fr = [rel(x, y) for x, y in zip(res["fp32"][1], res["fp64"][1])]
    print(f"  fields after 100 steps, rel-L2 (u, v, w, phi): {', '.join('%.2e' % x for x in fr)}", flush=True)
    assert max(fr) <= 1e-5, fr


@pytest.mark.timeout(600)
def test_ensemble_full_configuration(g, engine):
    """The uneven ensemble at its stated shape: 32 x 64^3 + 16 x 96^3 + 8 x 128^3 FP32 members
    (theta ~ U[0.03, 0.08], Re ~ U[500, 2000] from each member's seed, Sc 0.7), one hipGraph per
    RK3 step, 200 steps.  Every member is compared with its own FP64 run after 100 and 200 steps
    (the precision form's bound, rel-L2 <= 1e-5, holds at both)."""
    cfgs = g.ensemble_configs()
    assert len(cfgs) == 56
    print(f"\n[ensemble {len(cfgs)} members fp32, 200 steps]", flush=True)
    ens = g.Ensemble(cfgs, engine)
    ens.init()
    ens.step(100)
    mid = [[ens.get_spectral(i, f) for f in range(4)] for i in range(len(cfgs))]
    t1 = time.perf_counter()
    ens.step(100)
    t2 = time.perf_counter()
    end = [[ens.get_spectral(i, f) for f in range(4)] for i in range(len(cfgs))]
    print(f"  steps 101-200: {100 / (t2 - t1):.1f} ensemble steps/s", flush=True)
    e100, e200 = [], []
    for i, c in enumerate(cfgs):
        c64 = g.Config(**{**c.__dict__, "precision": g.FP64})
        s = g.Solver(c64, engine).setup()
        s.step(100)
        e100.append(max(rel(mid[i][f], s.get_spectral(f)) for f in range(4)))
        s.step(100)
        e200.append(max(rel(end[i][f], s.get_spectral(f)) for f in range(4)))
        st32, st64 = ens.stats(i), s.stats(False, False)[0]
        assert st32["step"] == 200 and abs(st32["ke"] - st64["ke"]) <= 1e-4 * st64["ke"], i
        s.close()
        if i % 8 == 7:
            print(f"  members 0-{i} compared", flush=True)
    for n in (64, 96, 128):
        idx = [i for i, c in enumerate(cfgs) if c.n[0] == n]
        w1 = max(idx, key=lambda i: e100[i])
        w2 = max(idx, key=lambda i: e200[i])
        print(f"  {n}^3 ({len(idx)} members): worst rel-L2 vs FP64 after 100 steps {e100[w1]:.2e} (member {w1}, "
              f"Re {cfgs[w1].re:.0f}, theta {cfgs[w1].theta:.4f}); after 200 steps {e200[w2]:.2e} (member {w2})",
              flush=True)
    ens.close()
    assert max(e100) <= 1e-5, max(e100)
    assert max(e200) <= 1e-5, max(e200)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("member", [0, 32, 48])
def test_ensemble_member_fp64_vs_oracle(g, engine, oracle, member):
    """One member per grid of the ensemble: its FP64 form against the oracle after 20 steps."""
    cfgs = g.ensemble_configs()
    c64 = g.Config(**{**cfgs[member].__dict__, "precision": g.FP64})
    _fftw(oracle, g, True)
    try:
        a, b = g.Solver(c64, engine).setup(), g.Solver(c64, oracle).setup()
        a.step(20)
        b.step(20)
        fr = [rel(a.get_spectral(f), b.get_spectral(f)) for f in range(4)]
        print(f"\n  member {member} ({c64.n[0]}^3) FP64 vs oracle after 20 steps: {max(fr):.1e}", flush=True)
        assert max(fr) <= 1e-12, fr
        a.close()
        b.close()
    finally:
        _fftw(oracle, g, False)


def _downsample(a, n_small):
    """Spectral field [ky][kx][kz] of a large grid -> the modes of an n_small grid."""
    ny, nx, nzh = a.shape
    h = n_small // 2
    iy = np.r_[0:h, ny - h:ny]
    ix = np.r_[0:h, nx - h:nx]
    return a[np.ix_(iy, ix, np.arange(h + 1))]


_J1024 = {}


@pytest.mark.timeout(600)
@pytest.mark.parametrize("form", ["mfma", "valu"])
def test_jet1024_fp32_runs(g, engine, form):
    """1024^3 FP32 (Re 1e4, theta 0.01, Sc 1, noise on |m| <= 64, dt 1e-4), 3 steps, with the
    split-FP16 MFMA radix-16 stage and with the VALU radix stages; the fields are kept on the
    512^3 modes for the comparisons below."""
    steps = 3
    flags = g.FLAG_MFMA_R16 if form == "mfma" else 0
    print(f"\n[jet1024 fp32 {form}: {steps} steps]", flush=True)
    s = g.Solver(g.config("jet1024", precision=g.FP32, flags=flags), engine).setup()
    t0 = time.perf_counter()
    s.step(steps)
    dt = time.perf_counter() - t0
    st = s.stats(False, False)[0]
    _J1024[form] = [_downsample(s.get_spectral(f), 512) for f in range(4)]
    print(f"  {form}: {1000 * dt / steps:.0f} ms/step, KE {st['ke']:.10e}, divergence {st['divergence']:.1e}",
          flush=True)
    s.close()
    assert st["divergence"] <= 1e-6


@pytest.mark.timeout(600)
def test_jet1024_mfma_vs_valu_and_fp64_512(g, engine):
    """The MFMA radix-16 stage against the VALU radix stages (rel-L2 <= 1e-6), and both against
    the FP64 plain form at 512^3 with the same physical parameters, compared on the 2/3-rule modes
    of the 512^3 grid (resolution + precision; rel-L2 <= 1e-4)."""
    if set(_J1024) != {"mfma", "valu"}:
        pytest.fail("the 1024^3 runs did not complete")
    mv = [rel(_J1024["mfma"][f], _J1024["valu"][f]) for f in range(4)]
    print(f"\n  MFMA vs VALU rel-L2 (u, v, w, phi): {', '.join('%.1e' % x for x in mv)}", flush=True)
    assert max(mv) <= 1e-6, mv
    s = g.Solver(g.config("jet1024", n=(512,) * 3, precision=g.FP64), engine).setup()
    s.step(3)
    ref = [s.get_spectral(f) * 8.0 for f in range(4)]  # unnormalised transforms: 1024^3 / 512^3 = 8
    mask = s.mask().astype(bool)
    s.close()
    for name in ("mfma", "valu"):
        d = [rel(_J1024[name][f][mask], ref[f][mask]) for f in range(4)]
        print(f"  {name} 1024^3 vs FP64 512^3 on the kept 512^3 modes: {', '.join('%.1e' % x for x in d)}", flush=True)
        assert max(d) <= 1e-4, (name, d)
