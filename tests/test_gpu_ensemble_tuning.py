"""Batched ensemble (one hipGraph per RK3 step) and the autotuned pass variants, on the GPU.

Each ensemble member must equal the same realisation advanced alone by gnss_step (same kernels),
and every tunable variant must give the same numbers as the default variant."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_ensemble_members_match_single_runs(g, engine, oracle):
    cfgs = g.ensemble_configs(counts=((32, 3), (48, 2), (64, 2)), precision=g.FP64)
    print(f"\n[ensemble {len(cfgs)} members]", flush=True)
    ens = g.Ensemble(cfgs, engine)
    ens.init()
    ens.step(4)
    for i, c in enumerate(cfgs):
        s = g.Solver(c, engine).setup()
        s.step(4)
        for f in range(4):
            assert rel(ens.get_spectral(i, f), s.get_spectral(f)) < 1e-14, (i, f)
        st = ens.stats(i)
        assert st["step"] == 4 and abs(st["t"] - 4 * c.dt) < 1e-15
    # and one member against the oracle
    o = g.Solver(cfgs[5], oracle).setup()
    o.step(4)
    for f in range(4):
        assert rel(ens.get_spectral(5, f), o.get_spectral(f)) < 1e-12


def test_batched_ensemble_equals_unbatched(g, engine, monkeypatch, tmp_path):
    """Members of one grid share one batched engine (their states back to back, one launch per
    pass for all of them); with GNSS_ENSEMBLE_BATCH=0 every member has an engine of its own.  With
    the same pass variants (a table without the batched "<key>b" entries) the per-element
    arithmetic is the same, so fields and statistics agree bit for bit; with the shipped table the
    batched engines take their own variants, which agree to FP32 rounding.  A batched engine always
    runs the late-derivative x mode (xf = 4), so the plain table gives the stand-alone engines that
    mode too."""
    import json
    from conftest import PKG
    table = json.load(open(f"{PKG}/tuning/gnss_tuning.json"))
    plain = tmp_path / "plain.json"
    plain.write_text("{\n" + ",\n".join(f'"{k}": {json.dumps(v if k.endswith("c") else dict(v, xf=4))}'
                                          for k, v in table.items() if not k.endswith("b")) + "\n}\n")
    cfgs = g.ensemble_configs(counts=((64, 5), (96, 3), (128, 2)))

    def run(batch, tuning):
        monkeypatch.setenv("GNSS_ENSEMBLE_BATCH", batch)
        if tuning:
            monkeypatch.setenv("GNSS_TUNING", tuning)
        else:
            monkeypatch.delenv("GNSS_TUNING", raising=False)
        print(f"\n[ensemble GNSS_ENSEMBLE_BATCH={batch} table={tuning or 'shipped'}, {len(cfgs)} members]",
              flush=True)
        ens = g.Ensemble(cfgs, engine)
        ens.init()
        ens.step(3)
        out = ([[ens.get_spectral(i, f) for f in range(4)] for i in range(len(cfgs))],
               [ens.stats(i) for i in range(len(cfgs))])
        ens.close()
        return out

    (fb, sb), (fu, su) = run("1", str(plain)), run("0", str(plain))
    for i in range(len(cfgs)):
        for f in range(4):
            assert np.array_equal(fb[i][f], fu[i][f]), (i, f)
        for k in ("ke", "epsilon", "cfl", "divergence", "t", "step"):
            assert sb[i][k] == su[i][k], (i, k)
    fs, _ = run("1", None)
    for i in range(len(cfgs)):
        for f in range(4):
            assert rel(fs[i][f], fu[i][f]) <= 2e-6, (i, f)


def test_ensemble_mixed_members_grouping(g, engine):
    """An ensemble mixing precisions, grids, scalar counts and a colloid member: members of one grid
    are batched, the colloid member stays on an engine of its own, and every member equals its
    stand-alone run (FP64 1e-13, FP32 2e-6: batched engines may take other pass variants)."""
    f32 = g.ensemble_configs(counts=((64, 2),))
    f64 = g.ensemble_configs(counts=((32, 2),), precision=g.FP64, seed=7)
    two = [g.Config(**{**c.__dict__, "sc": (0.7, 2.0), "seed": c.seed + 100}) for c in f32]
    col = [g.with_colloid(f32[0], alpha=0.02, theta_c=0.08)]
    cfgs = [f32[0], f64[0], two[0], col[0], f32[1], f64[1], two[1]]
    print(f"\n[ensemble of {len(cfgs)} mixed members]", flush=True)
    ens = g.Ensemble(cfgs, engine)
    ens.init()
    ens.step(3)
    for i, c in enumerate(cfgs):
        s = g.Solver(c, engine).setup()
        s.step(3)
        tol = 1e-13 if c.precision == g.FP64 else 2e-6
        for f in range(3 + len(c.sc)):
            assert rel(ens.get_spectral(i, f), s.get_spectral(f)) <= tol, (i, f)
        st = ens.stats(i)
        assert st["step"] == 3
        s.close()
    ens.close()


def test_ensemble_fp32_uneven(g, engine):
    cfgs = g.ensemble_configs(counts=((64, 4), (96, 2), (128, 1)))
    ens = g.Ensemble(cfgs, engine)
    ens.init()
    ens.step(3)
    for i in (0, 4, 6):
        s = g.Solver(cfgs[i], engine).setup()
        s.step(3)
        assert rel(ens.get_spectral(i, 0), s.get_spectral(0)) < 1e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tuning_variants_agree(g, engine, prec):
    p = g.FP64 if prec == "fp64" else g.FP32
    cfg = g.Config(n=(96, 64, 128), re=500.0, sc=(0.7,), theta=0.08, noise_kmax=8.0, noise_amp=0.2, dt=1e-3,
                   precision=p)
    specs = ["cx=4 cy=4 zf=0 xf=0 cp=0", "cx=8 cy=8 zf=0 xf=1 cp=1 xc=8", "cx=4 cy=8 zf=0 xf=1 cp=0 xc=4",
             "cx=8 cy=4 zf=3 xf=0 cp=1 zp=1", "cx=4 cy=8 xf=2 xc=4 cb=1 cp=1", "cx=8 cy=8 xf=3 xc=8 cb=1", "cx=8 cy=8 xf=0 cp=1 cb=1",
             "cx=8 cy=8 zf=6 xf=3 cb=1", "cx=8 cy=8 zf=7 xf=2 cb=1", "cx=8 cy=8 zf=8 xf=0 cp=1 cb=1"] if prec == "fp64" else \
        ["cx=8 cy=8 zf=0 xf=0 cp=0", "cx=16 cy=16 zf=0 xf=1 cp=1 xc=16", "cx=16 cy=8 zf=1 xf=1 cp=0 xc=8",
         "cx=8 cy=16 zf=1 xf=0 cp=1 zp=1", "cx=16 cy=16 zf=2 xf=0 cp=1", "cx=16 cy=16 zf=3 xf=1 cp=1",
         "cx=16 cy=16 zf=4 xf=0 cp=0", "cx=16 cy=8 zf=3 xf=2 xc=8", "cx=8 cy=16 xf=2 xc=16", "cx=16 cy=16 zf=3 xf=3 xc=8", "cx=16 cy=16 zf=5 xf=2 xc=16", "cx=16 cy=16 zf=3 xf=2 xc=16 cp=1 cb=1",
         "cx=8 cy=16 xf=0 cp=1 cb=1", "cx=16 cy=16 zf=6 xf=2 xc=16 cp=1 zp=1 cb=1",
         "cx=16 cy=16 zf=7 xf=3 xc=16 cb=1"]
    ref = None
    for spec in specs:
        s = g.Solver(cfg, engine).setup()
        s.set_tuning(spec)
        s.step(2)
        out = [s.get_spectral(f) for f in range(4)]
        if ref is None:
            ref = out
        else:
            tol = 1e-13 if prec == "fp64" else 2e-6
            for f in range(4):
                assert rel(out[f], ref[f]) < tol, (spec, f)


def test_autotune_reports_and_applies(g, engine):
    s = g.Solver(g.Config(n=(64, 64, 64), precision=g.FP32), engine).setup()
    s.step(1)
    before = [s.get_spectral(f) for f in range(4)]
    rep = s.autotune(3)
    for f in range(4):  # the timed substages run on a snapshot that is restored
        assert s.get_spectral(f).tobytes() == before[f].tobytes()
    assert "choice" in rep and "candidates" in rep
    assert any(k.startswith("cx=") for k in rep["candidates"])
    assert "cb=0 xf=0" in rep["candidates"] and "cp=1" in rep["candidates"] and "zp=1" in rep["candidates"]
    assert {"cb=0 xf=2 xc=8", "cb=1 xf=2 xc=16", "cb=0 xf=3 xc=8", "cb=1 xf=3 xc=16", "cb=1 xf=0"} <= set(rep["candidates"])
    assert {"cb=0 xf=1 xc=8", "cb=1 xf=1 xc=16"} <= set(rep["candidates"])
    assert {"cb=1 xf=4 xc=16 lp=0", "cb=1 xf=4 xc=16 lp=1", "cb=0 xf=5 xc=8 lp=1"} <= set(rep["candidates"])
    assert s.autotune(0)["choice"] == rep["choice"]
    with pytest.raises(g.GnssError):
        s.set_tuning("cx=3")
    with pytest.raises(g.GnssError):
        s.set_tuning("xc=4")  # FP32 fused x kernels are built for tiles 8 and 16


def test_bench_calls_report_rates(g, engine):
    """The measurement entry points: a plain copy of one spectral field and a 3D FFT pair."""
    s = g.Solver(g.Config(n=(128, 96, 64), sc=(0.7,), precision=g.FP32), engine).setup()
    ms, nbytes = s.bench_copy(3)
    kzs = -(-(64 // 2 + 1) // 16) * 16
    assert nbytes == 2 * 128 * 96 * kzs * 8 and ms > 0
    f, i = s.bench_fft(2)
    assert f > 0 and i > 0
    with pytest.raises(g.GnssError):
        s.bench_copy(0)


def test_ensemble_validates_members(g, engine):
    """gnss_ensemble_create checks members as gnss_create does, and refuses the unfused step (its
    physical buffers would be allocated inside the captured step)."""
    base = g.ensemble_configs(counts=((32, 2),), precision=g.FP64)
    for bad, code in [(dict(sc=(0.0,)), 1), (dict(precision=7), 1), (dict(flags=g.FLAG_UNFUSED), 6)]:
        cfgs = [base[0], g.Config(**{**base[1].__dict__, **bad})]
        with pytest.raises(g.GnssError) as e:
            g.Ensemble(cfgs, engine)
        assert e.value.code == code, bad


def test_ensemble_physical_field_is_the_members(g, engine):
    """GNSS_PHYSICAL through the ensemble is the selected member's current state transformed, not
    whatever another member's call left in the batched engine's shared physical buffers."""
    cfgs = g.ensemble_configs(counts=((32, 3),), precision=g.FP64)
    ens = g.Ensemble(cfgs, engine)
    ens.init()
    ens.step(2)
    ens.stats(2)  # leaves member 2's physical fields in the shared buffers
    for i in (0, 1, 2, 0):
        s = g.Solver(cfgs[i], engine).setup()
        s.step(2)
        for f in range(4):
            s.inverse(f)
            a, b = ens.get_physical(i, f), s.get_physical(f)
            assert rel(a, b) <= 1e-14, (i, f)
        s.close()
    ens.close()


def test_ensemble_autotune_tunes_batched_engines(g, engine):
    """gnss_ensemble_autotune times the batched engines' variants (their "<key>b" entries), keeps
    the states, captures the step again with the choices, and the members still equal their
    stand-alone runs."""
    cfgs = g.ensemble_configs(counts=((32, 4), (48, 3), (64, 1)))
    ens = g.Ensemble(cfgs, engine)
    ens.init()
    ens.step(1)
    before = [ens.get_spectral(i, 0) for i in range(len(cfgs))]
    reps = ens.autotune(1)
    assert sorted(r["key"] for r in reps) == ["f32 32 32 32 1b", "f32 48 48 48 1b"]
    for r in reps:
        assert {"zf=7", "xc=16 cp=1 cb=1 lp=1", "xc=8 cp=0 cb=0 lp=2"} <= set(r["candidates"]), r["key"]
        assert r["members"] == (4 if r["key"].startswith("f32 32") else 3)
    for i in range(len(cfgs)):
        assert ens.get_spectral(i, 0).tobytes() == before[i].tobytes(), i
    ens.step(2)
    for i in (0, 3, 4, 7):
        s = g.Solver(cfgs[i], engine).setup()
        s.step(3)
        for f in range(4):
            assert rel(ens.get_spectral(i, f), s.get_spectral(f)) <= 2e-6, (i, f)
        s.close()
    ens.close()
