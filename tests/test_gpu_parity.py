"""Engine (libgnss_hip.so, MI355X) against the FP64 oracle on the same inputs.

Bar: FP64 transforms <= 1e-13 relative; integer wavenumbers/masks bit-exact; initial condition
identical to the last bit; FP64 steps rel-L2 <= 1e-12 on u, v, w, phi."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


SIZES = [(16, 16, 16), (24, 32, 48), (64, 64, 64), (96, 96, 96), (128, 64, 32), (32, 384, 16)]


@pytest.mark.parametrize("n", SIZES, ids=lambda n: "x".join(map(str, n)))
def test_forward_inverse_fp64(g, engine, oracle, n):
    print(f"\n[fft fp64 {n}]", flush=True)
    cfg = g.Config(n=n, sc=())
    a, b = g.Solver(cfg, engine), g.Solver(cfg, oracle)
    x = np.random.default_rng(sum(n)).standard_normal(n)
    for s in (a, b):
        s.set_physical(0, x)
        s.forward(0)
    fa, fb = a.get_spectral(0), b.get_spectral(0)
    assert rel(fa, fb) < 1e-13
    ref = np.fft.rfftn(x).transpose(1, 0, 2)
    assert rel(fa, ref) < 1e-13
    for s in (a, b):
        s.inverse(0)
    assert np.abs(a.get_physical(0) - x).max() < 1e-12
    assert rel(a.get_physical(0), b.get_physical(0)) < 1e-13


@pytest.mark.parametrize("n", [(64, 64, 64), (96, 96, 96), (32, 384, 16)], ids=lambda n: "x".join(map(str, n)))
def test_forward_inverse_fp32(g, engine, oracle, n):
    print(f"\n[fft fp32 {n}]", flush=True)
    a = g.Solver(g.Config(n=n, sc=(), precision=g.FP32), engine)
    x = np.random.default_rng(7).standard_normal(n)
    a.set_physical(0, x)
    a.forward(0)
    ref = np.fft.rfftn(x).transpose(1, 0, 2)
    assert rel(a.get_spectral(0), ref) < 2e-6
    a.inverse(0)
    assert rel(a.get_physical(0), x) < 2e-6


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n", [(16, 16, 16), (96, 64, 128), (64, 32, 512), (32, 16, 1024)],
                         ids=lambda n: "x".join(map(str, n)))
def test_pipelined_z_passes_identical(g, engine, n, prec):
    """The persistent, pipelined z passes (zp=1) do the same arithmetic per row as the one-block-
    per-row-group passes: forward and inverse transforms agree to the last bit."""
    print(f"\n[zp {n} {prec}]", flush=True)
    x = np.random.default_rng(11).standard_normal(n)
    out = []
    for zp in (0, 1):
        a = g.Solver(g.Config(n=n, sc=(), precision=prec), engine)
        a.set_tuning(f"zp={zp}")
        a.set_physical(0, x)
        a.forward(0)
        spec = a.get_spectral(0)
        a.inverse(0)
        out.append((spec, a.get_physical(0)))
    assert out[0][0].tobytes() == out[1][0].tobytes()
    assert out[0][1].tobytes() == out[1][1].tobytes()
    assert rel(out[1][0], np.fft.rfftn(x).transpose(1, 0, 2)) < (1e-13 if prec == 0 else 2e-6)


@pytest.mark.parametrize("n", [(64, 64, 64), (24, 96, 48)], ids=lambda n: "x".join(map(str, n)))
def test_wavenumbers_and_mask_bit_exact(g, engine, oracle, n):
    L = (2 * np.pi, 3 * np.pi, 2 * np.pi)
    cfg = g.Config(n=n, L=L)
    a, b = g.Solver(cfg, engine), g.Solver(cfg, oracle)
    a.init_wavenumbers()
    b.init_wavenumbers()
    for ka, kb in zip(a.wavenumbers(), b.wavenumbers()):
        assert ka.tobytes() == kb.tobytes()
    assert a.mask().tobytes() == b.mask().tobytes()


def test_initial_condition_identical(g, engine, oracle):
    cfg = g.config("jet64")
    a, b = g.Solver(cfg, engine).setup(), g.Solver(cfg, oracle).setup()
    for f in range(4):
        assert a.get_spectral(f).tobytes() == b.get_spectral(f).tobytes(), f"field {f}"


@pytest.mark.parametrize("nsteps", [1, 5])
def test_step_fp64_32(g, engine, oracle, nsteps):
    print(f"\n[step fp64 32^3 x{nsteps}]", flush=True)
    cfg = g.Config(n=(32, 32, 32), re=300.0, sc=(0.7, 2.0), theta=0.1, noise_kmax=6.0, noise_amp=0.2, dt=2e-3)
    a, b = g.Solver(cfg, engine).setup(), g.Solver(cfg, oracle).setup()
    a.step(nsteps)
    b.step(nsteps)
    for f in range(5):
        assert rel(a.get_spectral(f), b.get_spectral(f)) < 1e-12, f


def test_jet64_100_steps_fp64(g, engine, oracle):
    """Plain FP64 temporal jet at 64^3, 100 steps: rel L2 <= 1e-12 on u, v, w, phi."""
    print("\n[jet64 fp64 x100]", flush=True)
    cfg = g.config("jet64")
    a, b = g.Solver(cfg, engine).setup(), g.Solver(cfg, oracle).setup()
    for chunk in range(5):
        a.step(20)
        b.step(20)
        errs = [rel(a.get_spectral(f), b.get_spectral(f)) for f in range(4)]
        print(f"  after {20 * (chunk + 1)} steps: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < 1e-12
    sa, ya, pa = a.stats()
    sb, yb, pb = b.stats()
    for k in ["ke", "epsilon", "enstrophy", "lint", "cfl", "umax"]:
        assert abs(sa[k] - sb[k]) <= 1e-10 * abs(sb[k]), k
    assert abs(sa["chi"][0] - sb["chi"][0]) <= 1e-10 * sb["chi"][0]
    assert np.abs(ya - yb).max() <= 1e-10 * np.abs(yb).max()
    assert np.abs(pa - pb).max() <= 1e-10 * np.abs(pb).max()


def test_stats_fp64(g, engine, oracle):
    cfg = g.Config(n=(48, 32, 64), sc=(0.7,), noise_kmax=5.0, noise_amp=0.1)
    a, b = g.Solver(cfg, engine).setup(), g.Solver(cfg, oracle).setup()
    sa, ya, pa = a.stats()
    sb, yb, pb = b.stats()
    for k in ["ke", "vrms", "epsilon", "enstrophy", "lint", "lambda", "eta", "cfl", "umax"]:
        assert abs(sa[k] - sb[k]) <= 1e-12 * abs(sb[k]), (k, sa[k], sb[k])
    assert abs(sa["divergence"]) < 1e-13
    assert np.abs(ya - yb).max() <= 1e-12 * np.abs(yb).max()
    assert np.abs(pa - pb).max() <= 1e-12 * np.abs(pb).max()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fused_matches_unfused(g, engine, prec):
    """The fused z kernel (C2R + products + R2C in registers) equals the separate-pass step."""
    p = g.FP64 if prec == "fp64" else g.FP32
    base = dict(n=(64, 48, 96), re=400.0, sc=(0.7, 2.0), theta=0.08, noise_kmax=8.0, noise_amp=0.2, dt=1e-3,
                precision=p)
    a = g.Solver(g.Config(**base), engine).setup()
    b = g.Solver(g.Config(**base, flags=g.FLAG_UNFUSED), engine).setup()
    a.step(3)
    b.step(3)
    tol = 1e-13 if prec == "fp64" else 1e-5
    for f in range(5):
        assert rel(a.get_spectral(f), b.get_spectral(f)) < tol, f


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nz", [16, 24, 48, 96, 192, 384, 512, 1024])
def test_fused_register_twiddles(g, engine, prec, nz):
    """Fused z variant 6 reads each thread's twiddles once and keeps them in registers (the packing
    factors of registers e >= E/2 taken as -i times those of e - E/2); variant 3 reads them from
    the tables inside every radix stage; variant 7 also moves the inverse transforms' 1/N onto the
    forward ones as 1/N^2, and the FP64 variant 8 pads its LDS rows differently.  All must give the
    same step, for every z length the engine supports, with a colloid so that every branch of the
    kernel runs."""
    cfg = g.with_colloid(g.Config(n=(32, 16, nz), re=400.0, sc=(0.7, 2.0), theta=0.08, noise_kmax=6.0,
                                  noise_amp=0.2, dt=1e-3, precision=prec), alpha=0.02)
    print(f"\n[register twiddles nz={nz} prec={prec}]", flush=True)
    out = []
    for zf in ((3, 6, 7, 8) if prec == 0 else (3, 6, 7)):
        a = g.Solver(cfg, engine).setup()
        a.set_tuning(f"zf={zf}")
        a.step(2)
        out.append([a.get_spectral(f) for f in range(5)])
        a.close()
    for o in out[1:]:
        for f in range(5):
            assert rel(o[f], out[0][f]) <= (1e-13 if prec == 0 else 2e-6), (nz, f)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nz", [16, 24, 48, 96, 192, 384, 512])
def test_fused_pair_rows_match(g, engine, prec, nz):
    """Fused z variant 9 transforms two real rows per complex transform of the full row length
    (z = a + i b; the half spectra packed with their Hermitian partners by lane shuffles, the
    products split again after the forward transform): the same step as variant 7, to rounding
    (FP64 1e-13, FP32 2e-6), in the late-derivative and the work-field x modes, with zero, one
    and two scalars (an odd or even number of product rows), and refused with a colloid."""
    print(f"\n[pair rows nz={nz} prec={prec}]", flush=True)
    for sc in ((), (0.7,), (0.7, 2.0)):
        for xf in (2, 5):
            cfg = g.Config(n=(32, 16, nz), re=400.0, sc=sc, theta=0.08, noise_kmax=6.0, noise_amp=0.2, dt=1e-3,
                           precision=prec)
            out = []
            for zf in (7, 9):
                a = g.Solver(cfg, engine).setup()
                a.set_tuning(f"zf={zf} xf={xf}")
                a.step(2)
                out.append([a.get_spectral(f) for f in range(3 + len(sc))])
                a.close()
            for f in range(3 + len(sc)):
                assert rel(out[1][f], out[0][f]) <= (1e-13 if prec == 0 else 2e-6), (sc, xf, f)
    c = g.Solver(g.with_colloid(g.Config(n=(32, 16, nz), sc=(0.7, 2.0), precision=prec)), engine)
    with pytest.raises(g.GnssError):
        c.set_tuning("zf=9")


@pytest.mark.parametrize("prec", [0, 1])
def test_per_field_x_inverse_matches(g, engine, prec):
    """x modes 2 and 3 form each work field inside its x inverse pass (u, omega, grad phi, the
    colloid's c and lap Z) with the operations of the prologue (mode 0) or of the group-fused
    inverse (mode 1) before the same transform; the compiler may contract a product and a sum into
    one FMA differently in the two kernels, so agreement is to rounding (FP64 1e-13, FP32 2e-6)."""
    cfg = g.with_colloid(g.Config(n=(64, 48, 96), re=400.0, sc=(0.7, 2.0), theta=0.08, noise_kmax=8.0,
                                  noise_amp=0.2, dt=1e-3, precision=prec), alpha=0.02)
    for pair in (("xf=0", "xf=2"), ("xf=1", "xf=3"), ("xf=2 cp=1 cb=0", "xf=2 cp=1 cb=1"), ("xf=0 cp=1 cb=0", "xf=0 cp=1 cb=1")):
        out = []
        for spec in pair:
            a = g.Solver(cfg, engine).setup()
            a.set_tuning(spec)
            a.step(3)
            out.append([a.get_spectral(f) for f in range(6)])
        for f in range(6):
            assert rel(out[1][f], out[0][f]) <= (1e-13 if prec == 0 else 2e-6), (pair, f)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,sc", [((64, 48, 96), (0.7, 2.0)), ((96, 64, 128), (1.0,)), ((32, 24, 48), ()),
                                  ((48, 32, 16), (0.5, 1.0, 2.0, 4.0))], ids=lambda v: "x".join(map(str, v)))
def test_late_derivative_modes_match(g, engine, prec, n, sc):
    """x modes 4 and 5 apply each derivative in the pass of its own axis (the x pass reads u, v, w,
    phi once; the y pass forms w_y, phi_y and omega_z = v_x - i k_y u; the fused z pass forms
    omega_x, omega_y and phi_z with i k_z as it loads).  Same step as modes 2 and 3, which form
    every work field from the state in the x pass, to rounding (FP64 1e-13, FP32 2e-6); with one
    launch per pass (cb = 1) and per field, with the two outputs of an entry in one launch (lp = 1,
    together; lp = 2, one after the other; lp = 3, together with the next tile loaded ahead) or
    split, with each fused z variant and with the y passes' own tile (yc)."""
    cfg = g.Config(n=n, re=400.0, sc=sc, theta=0.08, noise_kmax=8.0, noise_amp=0.2, dt=1e-3, precision=prec)
    nf = 3 + len(sc)
    zfs = (0, 3, 6, 7) if prec == 0 else (0, 1, 3, 6, 7)
    print(f"\n[late derivatives n={n} S={len(sc)} prec={prec}]", flush=True)
    for early, late in (("xf=2", "xf=4"), ("xf=3", "xf=5")):
        ref = g.Solver(cfg, engine).setup()
        ref.set_tuning(early + " cb=1")
        ref.step(3)
        want = [ref.get_spectral(f) for f in range(nf)]
        ref.close()
        ycs = ["cb=1 yc=%d" % c for c in ((4, 8) if prec == 0 else (8, 16))]
        for extra in ["cb=1 zf=%d" % z for z in zfs] + ["cb=0", "cb=1 lp=0", "cb=0 lp=0", "cb=1 lp=2", "cb=0 lp=2",
                                                        "cb=1 lp=3", "cb=0 lp=3"] + ycs:
            a = g.Solver(cfg, engine).setup()
            a.set_tuning(late + " " + extra)
            a.step(3)
            for f in range(nf):
                assert rel(a.get_spectral(f), want[f]) <= (1e-13 if prec == 0 else 2e-6), (late, extra, f)
            a.close()


def test_late_derivative_mode_matches_oracle(g, engine, oracle):
    """The 64^3 jet, 100 FP64 steps in x mode 4 against the oracle (rel-L2 <= 1e-12), and the mode
    is refused for a colloid, whose work fields need second derivatives."""
    cfg = g.config("jet64")
    a = g.Solver(cfg, engine).setup()
    a.set_tuning("xf=4")
    b = g.Solver(cfg, oracle).setup()
    a.step(100)
    b.step(100)
    for f in range(4):
        assert rel(a.get_spectral(f), b.get_spectral(f)) <= 1e-12, f
    c = g.Solver(g.with_colloid(g.config("jet64")), engine).setup()
    for spec in ("xf=4", "xf=5"):
        with pytest.raises(g.GnssError):
            c.set_tuning(spec)


def test_fp32_error_vs_fp64(g, engine):
    """FP32 form against the FP64 plain form on the same seed (64^3 jet, 100 steps)."""
    a = g.Solver(g.config("jet64", precision=g.FP32), engine).setup()
    b = g.Solver(g.config("jet64"), engine).setup()
    a.step(100)
    b.step(100)
    for f in range(4):
        assert rel(a.get_spectral(f), b.get_spectral(f)) < 1e-5, f
    sa, ya, _ = a.stats()
    sb, yb, _ = b.stats()
    for k in ["ke", "epsilon", "cfl"]:
        assert abs(sa[k] - sb[k]) < 1e-4 * abs(sb[k]), k
    assert abs(sa["chi"][0] - sb["chi"][0]) < 1e-4 * sb["chi"][0]
    assert sa["divergence"] < 1e-6


@pytest.mark.parametrize("n", [(256, 512, 16), (1024, 256, 16)], ids=lambda n: "x".join(map(str, n)))
def test_mfma_radix16_transform(g, engine, n):
    """FP32 transform with the split-FP16 MFMA radix-16 stage vs numpy and vs the VALU path."""
    print(f"\n[mfma fft {n}]", flush=True)
    x = np.random.default_rng(11).standard_normal(n)
    ref = np.fft.rfftn(x).transpose(1, 0, 2)
    out, names = {}, {}
    for flags in (0, g.FLAG_MFMA_R16):
        a = g.Solver(g.Config(n=n, sc=(), precision=g.FP32, flags=flags), engine)
        a.timing(True)
        a.set_physical(0, x)
        a.forward(0)
        out[flags] = a.get_spectral(0)
        a.inverse(0)
        assert rel(a.get_physical(0), x) < 3e-6
        names[flags] = set(a.timing_report())
    assert rel(out[g.FLAG_MFMA_R16], ref) < 3e-6
    assert rel(out[g.FLAG_MFMA_R16], out[0]) < 3e-6
    # the MFMA stage really ran: its passes are timed under their own names and its split-FP16
    # products round differently from the VALU butterflies
    assert {"fft_x_fwd_mfma", "fft_x_inv_mfma"} <= names[g.FLAG_MFMA_R16], names
    assert not any("mfma" in k for k in names[0])
    assert out[g.FLAG_MFMA_R16].tobytes() != out[0].tobytes()


def test_mfma_step_matches_valu(g, engine):
    base = dict(n=(256, 256, 32), re=800.0, sc=(1.0,), theta=0.05, noise_kmax=12.0, dt=5e-4,
                precision=g.FP32)
    a = g.Solver(g.Config(**base), engine).setup()
    b = g.Solver(g.Config(**base, flags=g.FLAG_MFMA_R16), engine).setup()
    a.step(5)
    b.step(5)
    for f in range(4):
        assert rel(b.get_spectral(f), a.get_spectral(f)) < 1e-5


@pytest.mark.parametrize("n,L,sc", [((48, 96, 24), (2 * np.pi, 4 * np.pi, np.pi), (0.7, 2.0)),
                                    ((24, 48, 32), (np.pi, 2 * np.pi, 2 * np.pi), ()),
                                    ((384, 16, 16), (4 * np.pi, np.pi, np.pi), (1.0,))],
                         ids=["48x96x24-aniso-S2", "24x48x32-S0", "384x16x16-S1"])
def test_step_ragged_anisotropic_vs_oracle(g, engine, oracle, n, L, sc):
    """Grids with factors of 3 on the axes, boxes of unequal sides (so the integer dealias mask uses
    its lcm form and the spectrum its anisotropic shells), no scalar or two: 20 FP64 steps in the
    tuned late-derivative mode against the oracle (fields rel-L2 <= 1e-12; statistics, planar
    profiles and shell spectrum to 1e-11)."""
    cfg = g.Config(n=n, L=L, re=500.0, sc=sc, theta=0.1, noise_kmax=5.0, noise_amp=0.1, dt=1e-3)
    a, b = g.Solver(cfg, engine).setup(), g.Solver(cfg, oracle).setup()
    assert np.array_equal(a.mask(), b.mask())
    a.step(20)
    b.step(20)
    for f in range(3 + len(sc)):
        assert rel(a.get_spectral(f), b.get_spectral(f)) <= 1e-12, f
    sa, ya, pa = a.stats()
    sb, yb, pb = b.stats()
    for k in ("ke", "epsilon", "enstrophy", "lint", "cfl", "umax"):
        assert abs(sa[k] - sb[k]) <= 1e-11 * abs(sb[k]), k
    for s in range(len(sc)):
        assert abs(sa["chi"][s] - sb["chi"][s]) <= 1e-11 * abs(sb["chi"][s])
    assert np.abs(ya - yb).max() <= 1e-11 * np.abs(yb).max()
    assert np.abs(pa - pb).max() <= 1e-11 * np.abs(pb).max()
