"""Pins the FP64 oracle (oracle/gnss_oracle.c) before it is trusted as the engine's checker.

The reference ships no tests or vectors (it has no test directory, and nothing it holds pins its
FFT library, cuFFT), so every check here is a brute-force computation, a committed
library-generated vector, a closed form or an invariant — chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails:

  FFT            brute-force DFT (explicit exponentials) on tiny grids; numpy.fft.rfftn vectors
                 committed in tests/golden/fft_golden.npz; round trip; Parseval
  wavenumbers    (2 pi / L) * m with the Nyquist index mapped to -n/2, bit-exact
  mask           exact rational enumeration of sum_i (m_i / n_i)^2 <= 1/9
  step           single-mode exact viscous decay (integrating factor), scalar diffusion, 2D
                 Taylor-Green exact decay (projection of a pure-gradient u x w), third-order
                 convergence of RK3, inviscid energy and scalar-variance conservation,
                 divergence-free evolution, Galilean sign check of the scalar advection and of
                 the velocity advection (a translating shear wave), one-step tendency of a 3D
                 field whose u x w is not a gradient (each vorticity component, the cross
                 product order, u . grad phi per component); FFTW3-API backend = built-in FFT
  statistics     single-mode closed forms (KE, epsilon, chi, spectrum shell, profiles, CFL)
  initial cond.  Hermitian symmetry, divergence, band limit, noise rms, jet DFT, seed determinism
"""
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN

TWO_PI = 2 * np.pi


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def solver(g, oracle, **kw):
    return g.Solver(g.Config(**kw), oracle)


def to_abi(X):
    """numpy rfftn order [kx][ky][kz] -> ABI order [ky][kx][kz]."""
    return np.ascontiguousarray(X.transpose(1, 0, 2))


def grid(n, L=(TWO_PI,) * 3):
    xs = [np.arange(n[i]) * L[i] / n[i] for i in range(3)]
    return np.meshgrid(*xs, indexing="ij")


# ------------------------------------------------------------------------------------------------
# FFT
# ------------------------------------------------------------------------------------------------
def brute_rfftn(x):
    nx, ny, nz = x.shape
    out = np.zeros((nx, ny, nz // 2 + 1), dtype=complex)
    Ex = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    Ey = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    Ez = np.exp(-2j * np.pi * np.outer(np.arange(nz // 2 + 1), np.arange(nz)) / nz)
    for a in range(nx):
        for b in range(ny):
            for c in range(nz // 2 + 1):
                out[a, b, c] = np.sum(x * Ex[a][:, None, None] * Ey[b][None, :, None] * Ez[c][None, None, :])
    return out


@pytest.mark.parametrize("n", [(4, 6, 8), (6, 4, 12), (8, 8, 8), (12, 6, 4)])
def test_fft_brute_force(g, oracle, n):
    s = solver(g, oracle, n=n, sc=())
    x = np.random.default_rng(1).standard_normal(n)
    s.set_physical(0, x)
    s.forward(0)
    assert rel(s.get_spectral(0), to_abi(brute_rfftn(x))) < 1e-14


def test_fft_golden_vectors(g, oracle):
    d = np.load(f"{GOLDEN}/fft_golden.npz")
    i = 0
    while f"x{i}" in d:
        n = tuple(int(v) for v in d[f"shape{i}"])
        s = solver(g, oracle, n=n, sc=())
        s.set_physical(0, d[f"x{i}"])
        s.forward(0)
        assert rel(s.get_spectral(0), to_abi(d[f"X{i}"])) < 1e-14, n
        s.inverse(0)
        assert np.abs(s.get_physical(0) - d[f"x{i}"]).max() < 1e-13, n
        i += 1
    assert i >= 4


def test_fft_parseval_and_delta(g, oracle):
    n = (8, 12, 16)
    s = solver(g, oracle, n=n, sc=())
    x = np.zeros(n)
    x[3, 5, 7] = 1.0  # delta -> pure phase exp(-2 pi i (3a/8 + 5b/12 + 7c/16))
    s.set_physical(0, x)
    s.forward(0)
    X = s.get_spectral(0)
    kb, ka, kc = np.meshgrid(np.arange(12), np.arange(8), np.arange(9), indexing="ij")
    ph = (3 * ka * 6 + 5 * kb * 4 + 7 * kc * 3) % 48  # exact phase in units of 2 pi / 48
    expect = np.exp(-2j * np.pi * ph / 48)
    assert np.abs(X - expect).max() < 1e-14
    y = np.random.default_rng(3).standard_normal(n)
    s.set_physical(0, y)
    s.forward(0)
    Y = s.get_spectral(0)
    w = np.full(Y.shape[2], 2.0)
    w[0] = w[-1] = 1.0
    assert abs((w * np.abs(Y) ** 2).sum() / np.prod(n) - (y ** 2).sum()) < 1e-12 * (y ** 2).sum()


# ------------------------------------------------------------------------------------------------
# Wavenumbers and mask
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("n,L", [((16, 16, 16), (TWO_PI,) * 3), ((24, 36, 12), (TWO_PI, 3 * np.pi, TWO_PI))])
def test_wavenumbers_closed_form(g, oracle, n, L):
    s = solver(g, oracle, n=n, L=L)
    s.init_wavenumbers()
    kx, ky, kz = s.wavenumbers()
    for k, N, l, cnt in [(kx, n[0], L[0], n[0]), (ky, n[1], L[1], n[1]), (kz, n[2], L[2], n[2] // 2 + 1)]:
        m = [i if i < N // 2 else i - N for i in range(cnt)]
        expect = np.array([(2.0 * np.pi / l) * float(v) for v in m])
        assert k.tobytes() == expect.tobytes()
    assert kz[-1] == -(2.0 * np.pi / L[2]) * (n[2] // 2)  # Nyquist -> -n/2 (initialize.cu:685-688)


@pytest.mark.parametrize("n", [(16, 16, 16), (24, 12, 48), (18, 18, 18)])
def test_mask_exact_enumeration(g, oracle, n):
    s = solver(g, oracle, n=n)
    s.init_wavenumbers()
    m = s.mask()
    expect = np.zeros_like(m)
    for iy in range(n[1]):
        my = iy if iy < n[1] // 2 else iy - n[1]
        for ix in range(n[0]):
            mx = ix if ix < n[0] // 2 else ix - n[0]
            for iz in range(n[2] // 2 + 1):
                mz = iz if iz < n[2] // 2 else iz - n[2]
                r = Fraction(mx, n[0]) ** 2 + Fraction(my, n[1]) ** 2 + Fraction(mz, n[2]) ** 2
                expect[iy, ix, iz] = 1 if r <= Fraction(1, 9) else 0
    assert m.tobytes() == expect.tobytes()
    assert m[:, :, -1].max() == 0  # Nyquist plane always removed


# ------------------------------------------------------------------------------------------------
# Step
# ------------------------------------------------------------------------------------------------
def set_physical_state(s, fields):
    """fields: list of physical arrays [x][y][z] -> spectral state (exact DFT via the oracle)."""
    for f, a in enumerate(fields):
        s.set_physical(f, a)
        s.forward(f)


def test_single_mode_exact_viscous_decay(g, oracle):
    n, re, dt, steps = (16, 16, 16), 50.0, 0.01, 20
    s = solver(g, oracle, n=n, re=re, sc=(0.7,), dt=dt, noise_amp=0.0)
    s.init_wavenumbers()
    X, Y, Z = grid(n)
    u0 = np.sin(2 * Y + 3 * Z)          # k = (0, 2, 3), velocity along x: k . a = 0
    phi0 = np.cos(X + 2 * Z)             # scalar mode |k|^2 = 5, advected by u (u . grad phi != 0)
    set_physical_state(s, [u0, 0 * u0, 0 * u0, 0 * u0])
    s.step(steps)
    t = steps * dt
    for f in range(3):
        s.inverse(f)
    assert np.abs(s.get_physical(0) - u0 * np.exp(-13 * t / re)).max() < 1e-13
    assert np.abs(s.get_physical(1)).max() < 1e-14 and np.abs(s.get_physical(2)).max() < 1e-14
    # scalar at rest: pure diffusion with D = nu / Sc
    s2 = solver(g, oracle, n=n, re=re, sc=(0.7,), dt=dt, noise_amp=0.0)
    s2.init_wavenumbers()
    set_physical_state(s2, [0 * u0, 0 * u0, 0 * u0, phi0])
    s2.step(steps)
    s2.inverse(3)
    assert np.abs(s2.get_physical(3) - phi0 * np.exp(-5 * t / (re * 0.7))).max() < 1e-13


def test_taylor_green_2d_exact_decay(g, oracle):
    n, re, dt, steps = (16, 16, 8), 100.0, 0.01, 30
    s = solver(g, oracle, n=n, re=re, sc=(), dt=dt, noise_amp=0.0)
    s.init_wavenumbers()
    X, Y, Z = grid(n)
    u0, v0 = np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y)
    set_physical_state(s, [u0, v0, 0 * u0])
    s.step(steps)
    decay = np.exp(-2 * steps * dt / re)
    for f, ref in [(0, u0), (1, v0)]:
        s.inverse(f)
        assert np.abs(s.get_physical(f) - ref * decay).max() < 1e-13
    s.inverse(2)
    assert np.abs(s.get_physical(2)).max() < 1e-14


def test_scalar_advection_sign(g, oracle):
    """Uniform flow U in x translates phi: phi(x - U t).  A wrong sign of u.grad(phi) fails."""
    n, dt, steps, U = (16, 8, 8), 1e-3, 50, 0.7
    s = solver(g, oracle, n=n, re=1e12, sc=(1.0,), dt=dt, noise_amp=0.0)
    s.init_wavenumbers()
    X, Y, Z = grid(n)
    phi0 = np.cos(2 * X) + 0.5 * np.sin(3 * X + Y)
    set_physical_state(s, [U + 0 * X, 0 * X, 0 * X, phi0])
    s.step(steps)
    s.inverse(3)
    t = steps * dt
    expect = np.cos(2 * (X - U * t)) + 0.5 * np.sin(3 * (X - U * t) + Y)
    assert np.abs(s.get_physical(3) - expect).max() < 1e-9


def test_shear_wave_galilean_translation(g, oracle):
    """u = (U, eps sin kx, 0): omega = (0, 0, eps k cos kx) and the y component of u x omega is
    -U eps k cos kx, so v = eps exp(-nu k^2 t) sin(k (x - U t)) travels at +U.  With omega negated
    it would travel at -U; with omega_z formed from the wrong derivatives it would not move.  The
    x component (eps^2 k / 2) sin 2kx is a pure gradient and the projection removes it, so u stays
    uniform (solver.cu:68-114 vorticity, 139-164 cross product, 504-511 projection)."""
    n, re, dt, steps, U, eps, k = (16, 8, 8), 200.0, 1e-3, 60, 0.8, 0.05, 2
    s = solver(g, oracle, n=n, re=re, sc=(), dt=dt, noise_amp=0.0)
    s.init_wavenumbers()
    X, Y, Z = grid(n)
    set_physical_state(s, [U + 0 * X, eps * np.sin(k * X), 0 * X])
    s.step(steps)
    t = steps * dt
    for f in range(3):
        s.inverse(f)
    expect_v = eps * np.exp(-k * k * t / re) * np.sin(k * (X - U * t))
    assert np.abs(s.get_physical(1) - expect_v).max() < 1e-12
    assert np.abs(s.get_physical(0) - U).max() < 1e-13
    assert np.abs(s.get_physical(2)).max() < 1e-14


def test_velocity_and_scalar_tendency_3d(g, oracle):
    """One small step from u = (a sin z, b sin x, c sin y), phi = cos x + 2 cos y + 3 cos z.

    omega = curl u = (c cos y, a cos z, b cos x), and
    u x omega = (b^2 sin x cos x - c a sin y cos z,  c^2 sin y cos y - a b sin z cos x,
                 a^2 sin z cos z - b c sin x cos y).
    The squared terms are gradients (removed by the projection); the mixed terms are
    perpendicular to their wavevectors and survive, so with nu = 1/Re and |k| = 1
        du/dt = -c a sin y cos z - nu a sin z,   dv/dt = -a b sin z cos x - nu b sin x,
        dw/dt = -b c sin x cos y - nu c sin y,
        dphi/dt = -u . grad phi + D lap phi
                = a sin z sin x + 2 b sin x sin y + 3 c sin y sin z - D (cos x + 2 cos y + 3 cos z).
    u x omega is not a gradient here, so this pins the sign of every vorticity component, the
    order of the cross product and which velocity component multiplies which gradient (a sign
    flip or a swapped component changes the tendency by O(1), against the O(dt) truncation)."""
    n, re, sc, dt = (16, 16, 16), 10.0, 0.5, 1e-6
    a, b, c = 0.7, -1.1, 1.3
    nu, D = 1.0 / re, 1.0 / (re * sc)
    s = solver(g, oracle, n=n, re=re, sc=(sc,), dt=dt, noise_amp=0.0)
    s.init_wavenumbers()
    X, Y, Z = grid(n)
    q0 = [a * np.sin(Z), b * np.sin(X), c * np.sin(Y), np.cos(X) + 2 * np.cos(Y) + 3 * np.cos(Z)]
    set_physical_state(s, q0)
    s.step(1)
    rate = []
    for f in range(4):
        s.inverse(f)
        rate.append((s.get_physical(f) - q0[f]) / dt)
    expect = [
        -c * a * np.sin(Y) * np.cos(Z) - nu * a * np.sin(Z),
        -a * b * np.sin(Z) * np.cos(X) - nu * b * np.sin(X),
        -b * c * np.sin(X) * np.cos(Y) - nu * c * np.sin(Y),
        a * np.sin(Z) * np.sin(X) + 2 * b * np.sin(X) * np.sin(Y) + 3 * c * np.sin(Y) * np.sin(Z)
        - D * (np.cos(X) + 2 * np.cos(Y) + 3 * np.cos(Z)),
    ]
    for f in range(4):
        scale = np.abs(expect[f]).max()
        assert np.abs(rate[f] - expect[f]).max() < 1e-4 * scale, f


def test_fftw_backend_matches_builtin(g, oracle):
    """The FFTW3-API backend (used as the checker of the long 256^3 comparison) gives the
    built-in Stockham FFT's transforms and step results."""
    oracle.gnss_oracle_set_backend.argtypes = [g.ctypes.c_int]
    oracle.gnss_oracle_backend_name.restype = g.ctypes.c_char_p
    if oracle.gnss_oracle_set_backend(1) != 0:
        oracle.gnss_oracle_set_backend(0)
        pytest.skip("no FFTW3 API library (libfftw3 or MKL's libmkl_rt) on this host")
    try:
        assert oracle.gnss_oracle_backend_name().decode() != "builtin-stockham"
        n = (12, 16, 24)
        x = np.random.default_rng(7).standard_normal(n)
        s = solver(g, oracle, n=n, sc=())
        s.set_physical(0, x)
        s.forward(0)
        X_fftw = s.get_spectral(0)
        s.inverse(0)
        x_fftw = s.get_physical(0)
        cfg = dict(n=(16, 16, 16), re=300.0, sc=(0.7,), theta=0.1, noise_amp=0.1, noise_kmax=4.0, dt=2e-3)
        a = solver(g, oracle, **cfg).setup()
        a.step(5)
        step_fftw = [a.get_spectral(f) for f in range(4)]
    finally:
        oracle.gnss_oracle_set_backend(0)
    assert rel(X_fftw, to_abi(brute_rfftn(x))) < 1e-14
    assert np.abs(x_fftw - x).max() < 1e-13
    b = solver(g, oracle, **cfg).setup()
    b.step(5)
    for f in range(4):
        assert rel(step_fftw[f], b.get_spectral(f)) < 1e-13, f


def jet_solver(g, oracle, dt, **kw):
    base = dict(n=(16, 16, 16), re=200.0, sc=(0.7,), theta=0.15, noise_amp=0.3, noise_kmax=4.0, dt=dt)
    base.update(kw)
    return solver(g, oracle, **base).setup()


def test_rk3_third_order(g, oracle):
    T = 0.24
    finals = {}
    for dt in [0.04, 0.02, 0.01, 0.0025]:
        s = jet_solver(g, oracle, dt)
        s.step(int(round(T / dt)))
        finals[dt] = np.concatenate([s.get_spectral(f).ravel() for f in range(4)])
    ref = finals[0.0025]
    e = [rel(finals[dt], ref) for dt in [0.04, 0.02, 0.01]]
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 6.0 < r1 < 10.5 and 6.0 < r2 < 10.5, (e, r1, r2)


def test_inviscid_conservation_and_divergence(g, oracle):
    s = jet_solver(g, oracle, 1e-3, re=1e14, sc=(1.0,), noise_amp=0.5)
    st0, _, _ = s.stats(False, False)
    s.step(10)
    st1, _, _ = s.stats(False, False)
    assert abs(st1["ke"] - st0["ke"]) < 1e-10 * st0["ke"]
    assert abs(st1["scalar_var"][0] - st0["scalar_var"][0]) < 1e-10 * st0["scalar_var"][0]
    assert st1["divergence"] < 1e-14
    # with viscosity the energy must decay
    s2 = jet_solver(g, oracle, 1e-3, re=100.0)
    a, _, _ = s2.stats(False, False)
    s2.step(5)
    b, _, _ = s2.stats(False, False)
    assert b["ke"] < a["ke"]


def test_step_applies_the_mask(g, oracle):
    s = jet_solver(g, oracle, 1e-3)
    m = s.mask().astype(bool)
    u = s.get_spectral(0)
    u[~m] = 1.0  # garbage in removed modes: the step must drop it (deAlias, solver.cu:704)
    s.set_spectral(0, u)
    s.step(1)
    assert np.abs(s.get_spectral(0)[~m]).max() == 0.0


# ------------------------------------------------------------------------------------------------
# Statistics
# ------------------------------------------------------------------------------------------------
def test_statistics_single_mode_closed_forms(g, oracle):
    n, re, sc, dt, A, B = (16, 16, 16), 400.0, 0.5, 2e-3, 0.3, 0.2
    s = solver(g, oracle, n=n, re=re, sc=(sc,), dt=dt)
    s.init_wavenumbers()
    X, Y, Z = grid(n)
    u0 = A * np.cos(2 * Y)
    phi0 = B * np.cos(3 * Z) + 0.25
    set_physical_state(s, [u0, 0 * u0, 0 * u0, phi0])
    st, yp, sp = s.stats()
    nu = 1 / re
    assert abs(st["ke"] - A * A / 4) < 1e-14
    assert abs(st["vrms"] - np.sqrt(A * A / 6)) < 1e-14
    assert abs(st["epsilon"] - nu * 4 * A * A / 2) < 1e-15
    assert abs(st["enstrophy"] - 4 * A * A / 4) < 1e-14
    assert abs(st["lint"] - 3 * np.pi / 4 * (A * A / 2 / 4) / (A * A / 4)) < 1e-14
    assert abs(st["chi"][0] - 2 * (nu / sc) * 9 * B * B / 2) < 1e-15
    assert abs(st["scalar_var"][0] - B * B / 2) < 1e-14
    assert abs(st["cfl"] - dt * A * n[0] / TWO_PI) < 1e-14
    assert abs(st["umax"] - A) < 1e-14
    y = np.arange(n[1]) * TWO_PI / n[1]
    assert np.abs(yp[0] - A * np.cos(2 * y)).max() < 1e-14     # <u>(y)
    assert np.abs(yp[1:7]).max() < 1e-14                      # V, W, stresses
    assert np.abs(yp[7] - 0.25).max() < 1e-14                 # <phi>(y)
    assert np.abs(yp[8] - B * B / 2).max() < 1e-14            # <phi'phi'>
    assert np.abs(yp[9]).max() < 1e-14                        # <v'phi'>
    expect_u = np.zeros(sp.shape[1])
    expect_u[2] = A * A / 4
    assert np.abs(sp[0] - expect_u).max() < 1e-14
    expect_s = np.zeros(sp.shape[1])
    expect_s[3] = B * B / 4
    assert abs(sp[1][0] - 0.25 ** 2 / 2) < 1e-14 and abs(sp[1][3] - B * B / 4) < 1e-14
    assert abs(sp[0].sum() - st["ke"]) < 1e-14


def test_statistics_reynolds_stress(g, oracle):
    n = (16, 8, 16)
    s = solver(g, oracle, n=n, sc=(1.0,))
    s.init_wavenumbers()
    X, Y, Z = grid(n)
    u = 0.5 + 0.1 * np.cos(X) * (1 + np.cos(Y))
    v = 0.2 * np.cos(X) * (1 + np.cos(Y))
    set_physical_state(s, [u, v, 0 * u, v + 1.0])
    _, yp, _ = s.stats()
    f = 1 + np.cos(np.arange(8) * TWO_PI / 8)
    assert np.abs(yp[0] - 0.5).max() < 1e-14
    assert np.abs(yp[3] - 0.01 * f * f / 2).max() < 1e-14   # <u'u'>
    assert np.abs(yp[4] - 0.04 * f * f / 2).max() < 1e-14   # <v'v'>
    assert np.abs(yp[6] - 0.02 * f * f / 2).max() < 1e-14   # <u'v'>
    assert np.abs(yp[9] - 0.04 * f * f / 2).max() < 1e-14   # <v'phi'>


# ------------------------------------------------------------------------------------------------
# Initial condition
# ------------------------------------------------------------------------------------------------
def test_initial_condition_properties(g, oracle):
    cfg = dict(n=(32, 24, 32), re=1000.0, sc=(0.7,), theta=0.05, noise_amp=0.05, noise_kmax=6.0)
    s = solver(g, oracle, **cfg).setup()
    s0 = solver(g, oracle, **{**cfg, "noise_amp": 0.0}).setup()
    nx, ny, nz = cfg["n"]
    nn = nx * ny * nz
    st, _, _ = s.stats(False, False)
    assert st["divergence"] < 1e-15
    kx, ky, kz = s.wavenumbers()
    w = np.full(nz // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    energy = 0.0
    for f in range(3):
        d = s.get_spectral(f) - s0.get_spectral(f)
        my = np.array([i if i < ny // 2 else i - ny for i in range(ny)])[:, None, None]
        mx = np.array([i if i < nx // 2 else i - nx for i in range(nx)])[None, :, None]
        mz = np.arange(nz // 2 + 1)[None, None, :]
        m2 = mx ** 2 + my ** 2 + mz ** 2
        assert np.abs(d[(m2 > 36) | (m2 == 0)]).max() == 0.0   # band 1 <= |m| <= 6
        energy += (w * np.abs(d) ** 2).sum() / nn ** 2
        # Hermitian on kz = 0: X(-mx, -my) = conj X(mx, my)
        p = d[:, :, 0]
        pr = np.conj(p[(-np.arange(ny)) % ny][:, (-np.arange(nx)) % nx])
        assert np.abs(p - pr).max() < 1e-12 * np.abs(p).max()
    assert abs(np.sqrt(energy / 3) - 0.05) < 1e-14               # per-component rms
    # jet: exact DFT of the sampled profile along y, at kx = kz = 0, masked
    y = np.arange(ny) * TWO_PI / ny
    U = 0.5 * (1 + np.tanh((0.5 - np.abs(y - np.pi)) / (2 * 0.05)))
    J = nx * nz * np.array([np.sum(U * np.exp(-2j * np.pi * k * np.arange(ny) / ny)) for k in range(ny)])
    keep = s0.mask()[:, 0, 0].astype(bool)
    u0 = s0.get_spectral(0)
    assert np.abs(u0[keep, 0, 0] - J[keep]).max() < 1e-12 * np.abs(J).max()
    assert np.abs(u0[~keep, 0, 0]).max() == 0.0
    phi = s0.get_spectral(3)
    assert np.abs(phi - u0).max() == 0.0                        # same profile on the scalar
    assert np.abs(s0.get_spectral(1)).max() == 0.0


def test_initial_condition_seeded(g, oracle):
    a = solver(g, oracle, n=(16, 16, 16), noise_kmax=4.0, seed=1).setup()
    b = solver(g, oracle, n=(16, 16, 16), noise_kmax=4.0, seed=1).setup()
    c = solver(g, oracle, n=(16, 16, 16), noise_kmax=4.0, seed=2).setup()
    assert a.get_spectral(1).tobytes() == b.get_spectral(1).tobytes()
    assert rel(a.get_spectral(1), c.get_spectral(1)) > 0.5
