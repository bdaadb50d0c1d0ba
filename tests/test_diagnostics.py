"""Diagnostics: plane slices (writexyfields / save2Dfields, iofuncs.cu:187-327) and iso-surface
areas (surfaceArea_kernel_mgpu / calcSurfaceArea_mgpu, statistics.cu:1310-1492; |omega| from
VectorMagnitude, statistics.cu:2005-2037).

CPU tests pin the oracle:
  * a planar iso-surface F = cos x = iso crosses every (y, z) line twice; the indicator-gradient
    sum adds exactly 1/dx per crossing and line, so the area is exactly 2 Ly Lz;
  * a sphere |x - c| = r approaches 4 pi r^2 (first-order in dx);
  * |omega| of the shear layer u = sin y is |cos y| (slice), and |cos y| = 1/2 is four planes;
  * slices equal the matching planes of the physical field; the files have the reference's
    size header and [x][y] order, and the colloid file is named c.
GPU tests compare the engine with the oracle on an evolved jet.
"""
import os

import numpy as np
import pytest

TWO_PI = 2 * np.pi


def grid(n, L):
    xs = [np.arange(n[i]) * L[i] / n[i] for i in range(3)]
    return np.meshgrid(*xs, indexing="ij")


def state(g, lib, n, L, fields, sc=(1.0,), **kw):
    s = g.Solver(g.Config(n=n, L=L, sc=sc, noise_amp=0.0, **kw), lib)
    s.init_wavenumbers()
    for f, a in enumerate(fields):
        s.set_physical(f, a)
        s.forward(f)
    return s


@pytest.mark.parametrize("iso", [0.0, 0.5, -0.3])
def test_planar_iso_surface_is_exact(g, oracle, iso):
    n, L = (32, 24, 16), (TWO_PI, 3.0, 2.0)
    X, Y, Z = grid(n, L)
    s = state(g, oracle, n, L, [0 * X, 0 * X, 0 * X, np.cos(X)])
    assert abs(s.surface_area(3, iso) - 2 * L[1] * L[2]) < 1e-12 * L[1] * L[2]


def test_sphere_area_converges(g, oracle):
    errs = []
    for m in (32, 64):
        n, L = (m,) * 3, (TWO_PI,) * 3
        X, Y, Z = grid(n, L)
        r = 1.3
        F = np.sqrt((X - np.pi) ** 2 + (Y - np.pi) ** 2 + (Z - np.pi) ** 2)
        s = state(g, oracle, n, L, [0 * X, 0 * X, 0 * X, F])
        errs.append(abs(s.surface_area(3, r) / (4 * np.pi * r * r) - 1))
    assert errs[1] < 0.03 and errs[1] < errs[0]


def test_vorticity_magnitude_slice_and_area(g, oracle):
    n, L = (16, 32, 8), (TWO_PI,) * 3
    X, Y, Z = grid(n, L)
    s = state(g, oracle, n, L, [np.sin(Y), 0 * X, 0 * X, 0 * X])
    w = s.slice(g.FIELD_VORTICITY, 2, 3)                     # [x][y]
    assert np.abs(w - np.abs(np.cos(Y[:, :, 3]))).max() < 1e-13
    assert abs(s.surface_area(g.FIELD_VORTICITY, 0.5) - 4 * L[0] * L[2]) < 1e-11


def test_slices_match_physical_planes_and_files(g, oracle, tmp_path):
    n, L = (16, 12, 8), (TWO_PI,) * 3
    X, Y, Z = grid(n, L)
    rng = np.random.default_rng(5)
    fields = [rng.standard_normal(n) for _ in range(5)]
    cfg = dict(flags=g.FLAG_COLLOID, alpha=0.01)
    s = state(g, oracle, n, L, fields, sc=(0.7, 0.7), **cfg)
    for f in range(5):
        s.inverse(f)
        a = s.get_physical(f)
        assert np.abs(s.slice(f, 2, 5) - a[:, :, 5]).max() < 1e-13
        assert np.abs(s.slice(f, 1, 7) - a[:, 7, :]).max() < 1e-13
        assert np.abs(s.slice(f, 0, 2) - a[2, :, :]).max() < 1e-13
    s.save_slices(str(tmp_path), 40)
    names = ["u", "v", "w", "z", "c"]
    for f, nm in enumerate(names):
        raw = np.fromfile(os.path.join(tmp_path, f"{nm}_xy.40"), dtype=np.float64)
        assert raw[0] == n[0] * n[1] * 8
        assert np.abs(raw[1:].reshape(n[0], n[1]) - s.slice(f, 2, n[2] // 2)).max() == 0
    s.save_fields(str(tmp_path), 40)
    assert os.path.exists(os.path.join(tmp_path, "c.40")) and not os.path.exists(os.path.join(tmp_path, "z1.40"))


def test_diagnostic_argument_checks(g, oracle):
    n, L = (16, 16, 16), (TWO_PI,) * 3
    X, Y, Z = grid(n, L)
    s = state(g, oracle, n, L, [0 * X] * 4)
    for bad in [(4, 2, 0), (-2, 2, 0), (0, 3, 0), (0, 2, 16), (0, 1, -1)]:
        with pytest.raises(g.GnssError):
            s.slice(*bad)
    with pytest.raises(g.GnssError):
        s.surface_area(7, 0.0)


# ------------------------------------------------------------------------------------------------
# Engine parity (GPU)
# ------------------------------------------------------------------------------------------------
def evolved_jet(g, lib, prec=0):
    cfg = g.with_colloid(g.config("jet64", n=(64, 48, 32), L=(TWO_PI, 4.0, 3.0), precision=prec),
                         alpha=0.02)
    s = g.Solver(cfg, lib).setup()
    s.step(20)
    return s


@pytest.mark.gpu
def test_gpu_diagnostics_match_oracle(g, engine, oracle):
    a, b = evolved_jet(g, engine), evolved_jet(g, oracle)
    for f in (0, 3, 4, g.FIELD_VORTICITY):
        for axis, idx in ((2, 16), (1, 5), (0, 63)):
            sa, sb = a.slice(f, axis, idx), b.slice(f, axis, idx)
            assert np.linalg.norm(sa - sb) <= 1e-12 * np.linalg.norm(sb), (f, axis)
    for f, iso in ((3, 0.5), (4, 0.2), (0, 0.5), (g.FIELD_VORTICITY, 1.0)):
        aa, ab = a.surface_area(f, iso), b.surface_area(f, iso)
        assert ab > 0 and abs(aa - ab) <= 1e-12 * ab, (f, iso, aa, ab)


@pytest.mark.gpu
def test_gpu_diagnostics_fp32(g, engine):
    a, b = evolved_jet(g, engine, prec=1), evolved_jet(g, engine, prec=0)
    for f, iso in ((3, 0.5), (g.FIELD_VORTICITY, 1.0)):
        aa, ab = a.surface_area(f, iso), b.surface_area(f, iso)
        assert abs(aa - ab) <= 2e-3 * ab, (f, aa, ab)
    sa, sb = a.slice(3, 2, 16), b.slice(3, 2, 16)
    assert np.abs(sa - sb).max() <= 1e-5


@pytest.mark.gpu
def test_gpu_sphere_area(g, engine):
    n, L = (64,) * 3, (TWO_PI,) * 3
    X, Y, Z = grid(n, L)
    F = np.sqrt((X - np.pi) ** 2 + (Y - np.pi) ** 2 + (Z - np.pi) ** 2)
    s = state(g, engine, n, L, [0 * X, 0 * X, 0 * X, F])
    assert abs(s.surface_area(3, 1.3) / (4 * np.pi * 1.3 ** 2) - 1) < 0.03
