"""Colloid scalar (GNSS_FLAG_COLLOID): the last scalar c drifts with velocity alpha grad Z, Z = scalar 0.

Reference: colloidAdvectionKernel_mgpu / scalarAdvection (solver.cu:374-459) give the nonlinear
term u.grad c + alpha (lap Z c + grad Z . grad c); the colloid diffuses with nu / Sc_c
(solver.cu:603, 611) and starts from the jet profile of thickness theta_c (initialize.cu:178-194).

CPU tests pin the oracle by what the mathematics fixes:
  * with u = 0 and a frozen carrier Z (Sc_Z huge), c = exp(alpha Z / D_c) has zero flux
    alpha c grad Z - D_c grad c, so it is an exact steady state: a dropped lap Z term, a wrong
    sign of either drift term or a wrong diffusivity makes it move at O(alpha) per unit time;
  * a uniform flow U translates that steady state rigidly: c(x - U t) (pins u.grad c with drift);
  * alpha = 0 reduces the colloid to a plain scalar bit for bit;
  * theta_c sets the colloid's initial profile (compared with numpy's DFT of the tanh profile).
GPU tests compare the engine with the oracle (FP64, rel L2 <= 1e-12) and with its own FP64 form.
"""
import numpy as np
import pytest

TWO_PI = 2 * np.pi


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def grid(n):
    xs = [np.arange(n[i]) * TWO_PI / n[i] for i in range(3)]
    return np.meshgrid(*xs, indexing="ij")


def boltzmann_case(g, lib, n=(32, 32, 32), U=(0.0, 0.0, 0.0), alpha=0.005, re=100.0, sc_c=1.0,
                   dt=0.01, precision=0, alpha_run=None):
    """Frozen carrier Z = cos x + 0.5 cos(y + z) and c = exp(alpha Z / D_c) in uniform flow U,
    stepped with drift coefficient alpha_run (default alpha)."""
    cfg = g.Config(n=n, re=re, sc=(1e12, sc_c), dt=dt, noise_amp=0.0, flags=g.FLAG_COLLOID,
                   alpha=alpha if alpha_run is None else alpha_run, precision=precision)
    s = g.Solver(cfg, lib)
    s.init_wavenumbers()
    X, Y, Z = grid(n)
    d_c = 1.0 / (re * sc_c)

    def fields(shift):
        x, y, z = X - U[0] * shift, Y - U[1] * shift, Z - U[2] * shift
        zz = np.cos(x) + 0.5 * np.cos(y + z)
        return zz, np.exp(alpha * zz / d_c)

    z0, c0 = fields(0.0)
    for f, a in enumerate([U[0] + 0 * X, U[1] + 0 * X, U[2] + 0 * X, z0, c0]):
        s.set_physical(f, a)
        s.forward(f)
    s.dealias()
    return s, fields


def physical(s, f):
    s.inverse(f)
    return s.get_physical(f)


# ------------------------------------------------------------------------------------------------
# Oracle pins (CPU)
# ------------------------------------------------------------------------------------------------
def test_colloid_boltzmann_steady_state(g, oracle):
    s, fields = boltzmann_case(g, oracle)
    c_start = physical(s, 4)
    s.step(50)
    c_end = physical(s, 4)
    assert rel(c_end, c_start) < 1e-11
    # stepped without the drift, or with its sign flipped, the same state moves
    for alpha_run in (0.0, -0.005):
        s2, _ = boltzmann_case(g, oracle, alpha_run=alpha_run)
        s2.step(50)
        assert rel(physical(s2, 4), c_start) > 1e-4, alpha_run


def test_colloid_translation_in_uniform_flow(g, oracle):
    U, dt, steps = (0.6, -0.4, 0.3), 2e-3, 100
    s, fields = boltzmann_case(g, oracle, n=(32, 32, 32), U=U, dt=dt)
    s.step(steps)
    z_exp, c_exp = fields(steps * dt)
    assert np.abs(physical(s, 3) - z_exp).max() < 1e-8
    assert rel(physical(s, 4), c_exp) < 1e-8


def test_colloid_alpha_zero_is_a_plain_scalar(g, oracle):
    base = dict(n=(16, 16, 16), re=300.0, sc=(0.7, 1.3), theta=0.2, noise_amp=0.2, noise_kmax=4.0,
                dt=2e-3)
    a = g.Solver(g.Config(**base), oracle).setup()
    b = g.Solver(g.Config(**base, flags=g.FLAG_COLLOID, alpha=0.0), oracle).setup()
    a.step(5)
    b.step(5)
    for f in range(5):
        assert np.array_equal(a.get_spectral(f), b.get_spectral(f)), f


def test_colloid_initial_profile_theta_c(g, oracle):
    n, theta, theta_c = (16, 32, 8), 0.2, 0.35
    cfg = g.with_colloid(g.Config(n=n, sc=(0.7,), theta=theta, noise_amp=0.0), theta_c=theta_c)
    s = g.Solver(cfg, oracle).setup()
    y = np.arange(n[1]) * TWO_PI / n[1]
    mask = s.mask()
    for f, th in ((3, theta), (4, theta_c)):
        prof = 0.5 * (1.0 + np.tanh((0.5 - np.abs(y - np.pi)) / (2.0 * th)))
        spec = np.fft.fft(prof) * n[0] * n[2]
        got = s.get_spectral(f)[:, 0, 0]
        want = np.where(mask[:, 0, 0] > 0, spec, 0)
        assert np.abs(got - want).max() < 1e-12 * np.abs(spec).max()


def test_colloid_parameter_checks(g, oracle):
    with pytest.raises(g.GnssError):
        g.Solver(g.Config(n=(16,) * 3, sc=(0.7,), flags=g.FLAG_COLLOID, alpha=0.01), oracle)
    with pytest.raises(g.GnssError):
        g.Solver(g.Config(n=(16,) * 3, sc=(0.7, 0.7), flags=g.FLAG_COLLOID, alpha=float("nan")), oracle)


# ------------------------------------------------------------------------------------------------
# Engine parity (GPU)
# ------------------------------------------------------------------------------------------------
def colloid_jet(g, n=64, **over):
    return g.with_colloid(g.config("jet64", n=(n,) * 3, **over), alpha=0.02, theta_c=0.08)


@pytest.mark.gpu
def test_gpu_colloid_jet_matches_oracle(g, engine, oracle):
    cfg = colloid_jet(g)
    a = g.Solver(cfg, engine).setup()
    b = g.Solver(cfg, oracle).setup()
    for f in range(5):
        assert np.array_equal(a.get_spectral(f), b.get_spectral(f)), f
    a.step(100)
    b.step(100)
    for f in range(5):
        assert rel(a.get_spectral(f), b.get_spectral(f)) <= 1e-12, f


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, 2])
def test_gpu_colloid_two_plain_scalars_order(g, engine, oracle, flags):
    """Scalars 0, 1 plain and 2 the colloid: the fused kernel reorders the plain scalars."""
    cfg = g.with_colloid(g.config("jet64", n=(32,) * 3, sc=(0.7, 2.0), flags=flags), alpha=0.02)
    a = g.Solver(cfg, engine).setup()
    b = g.Solver(cfg, oracle).setup()
    a.step(10)
    b.step(10)
    for f in range(6):
        assert rel(a.get_spectral(f), b.get_spectral(f)) <= 1e-12, f


@pytest.mark.gpu
def test_gpu_colloid_boltzmann_steady_state(g, engine):
    s, _ = boltzmann_case(g, engine)
    c_start = physical(s, 4)
    s.step(50)
    assert rel(physical(s, 4), c_start) < 1e-11


@pytest.mark.gpu
def test_gpu_colloid_fp32_vs_fp64(g, engine):
    cfg64 = colloid_jet(g)
    cfg32 = g.Config(**{**cfg64.__dict__, "precision": g.FP32})
    a = g.Solver(cfg32, engine).setup()
    b = g.Solver(cfg64, engine).setup()
    a.step(100)
    b.step(100)
    for f in range(5):
        assert rel(a.get_spectral(f), b.get_spectral(f)) <= 1e-5, f


@pytest.mark.gpu
@pytest.mark.parametrize("prec,spec", [(0, "xf=0"), (0, "xf=0 cp=1"), (0, "xf=1 xc=4"), (0, "zf=3"), (1, "xf=1 zf=1"), (1, "zf=3"), (1, "zf=4"), (0, "xf=2"), (1, "xf=2 zf=3 xc=8"), (0, "xf=3"), (1, "xf=3 xc=16"), (0, "xf=2 cp=1 cb=1"), (1, "xf=2 cp=1 cb=1"), (1, "xf=0 cp=1 cb=1"),
                                       (1, "xf=1 zf=2 xc=8"), (1, "xf=0 cp=1"), (0, "zf=6 xf=3"), (1, "zf=6 xf=2 cp=1 cb=1"),
                                       (0, "zf=8 xf=2"), (1, "zf=7 xf=2 cp=1 cb=1")])
def test_gpu_colloid_tuning_variants(g, engine, prec, spec):
    """Every pass variant computes the same step (FP64: to rounding; FP32: to FP32 rounding)."""
    cfg = colloid_jet(g, precision=prec)
    a = g.Solver(cfg, engine).setup()
    b = g.Solver(cfg, engine).setup()
    a.set_tuning("xf=1")
    b.set_tuning(spec)
    a.step(5)
    b.step(5)
    tol = 1e-13 if prec == 0 else 1e-5
    for f in range(5):
        assert rel(b.get_spectral(f), a.get_spectral(f)) <= tol, (spec, f)
