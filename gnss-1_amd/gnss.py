"""Python host mirror of the C ABI in include/gnss.h (ctypes, no torch types).

The reference's host driver is C++ (main.cu:84-252); the C++ driver in driver/gnss_main.cpp keeps
that structure.  This module exposes the same entry points to Python for tests and bench.py:

    Solver(params)          ~ splitData + allocate_memory + plan*FFT   (main.cu:105-113)
    .init_wavenumbers()     ~ initializeWaveNumbers                    (initialize.cu:693-707)
    .init_jet()             ~ initializeJet + forward x5 + deAlias     (main.cu:139-156)
    .step(n)                ~ solver_ps                                (solver.cu:700-765)
    .stats()                ~ calcTurbStats_mgpu                       (statistics.cu:2039-2156)
    .get_field / set_field  ~ save3Dfields / importVelocity copies     (iofuncs.cu:329-535)
    .forward / .inverse     ~ forwardTransform / inverseTransform      (fftfuncs.cu:288-359)

`load()` loads the MI355X engine (libgnss_hip.so built in-tree).  It raises when the library is
missing: there is no CPU fallback.  Any other library implementing the same ABI can be passed
explicitly with `load(path)`.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_LIB = os.path.join(HERE, "lib", "libgnss_hip.so")

MAX_SCALARS = 4
FP64, FP32 = 0, 1
SPECTRAL, PHYSICAL = 0, 1
FLAG_MFMA_R16, FLAG_UNFUSED, FLAG_COLLOID = 1, 2, 4
FIELD_VORTICITY = -1

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_DEVICE", 3: "ERR_ALLOC", 4: "ERR_STATE",
          5: "ERR_NUMERIC", 6: "ERR_UNSUPPORTED"}


class GnssError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("re", ctypes.c_double),
        ("n_scalars", ctypes.c_int),
        ("sc", ctypes.c_double * MAX_SCALARS),
        ("theta", ctypes.c_double),
        ("jet_h", ctypes.c_double),
        ("noise_amp", ctypes.c_double),
        ("noise_kmax", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("dt", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("theta_c", ctypes.c_double),
    ]


class StatsOut(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double), ("step", ctypes.c_int64),
        ("ke", ctypes.c_double), ("vrms", ctypes.c_double), ("epsilon", ctypes.c_double),
        ("enstrophy", ctypes.c_double), ("lint", ctypes.c_double), ("lambda_", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("chi", ctypes.c_double * MAX_SCALARS),
        ("scalar_var", ctypes.c_double * MAX_SCALARS),
        ("cfl", ctypes.c_double), ("umax", ctypes.c_double), ("divergence", ctypes.c_double),
    ]

    def as_dict(self, n_scalars: int) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("chi", "scalar_var")}
        d["lambda"] = d.pop("lambda_")
        d["chi"] = [self.chi[i] for i in range(n_scalars)]
        d["scalar_var"] = [self.scalar_var[i] for i in range(n_scalars)]
        return d


class Sizes(ctypes.Structure):
    _fields_ = [("n_fields", ctypes.c_int), ("nzh", ctypes.c_int), ("n_prof", ctypes.c_int),
                ("n_bins", ctypes.c_int), ("n_physical", ctypes.c_int64),
                ("n_spectral", ctypes.c_int64)]


# Every entry point of include/gnss.h with its ctypes signature.
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "gnss_create": (_I, [ctypes.POINTER(Params), ctypes.POINTER(_P)]),
    "gnss_destroy": (_I, [_P]),
    "gnss_last_error": (ctypes.c_char_p, [_P]),
    "gnss_get_sizes": (_I, [_P, ctypes.POINTER(Sizes)]),
    "gnss_init_wavenumbers": (_I, [_P]),
    "gnss_get_wavenumbers": (_I, [_P, _D, _D, _D]),
    "gnss_get_mask": (_I, [_P, ctypes.POINTER(ctypes.c_uint8)]),
    "gnss_init_jet": (_I, [_P]),
    "gnss_forward": (_I, [_P, _I]),
    "gnss_inverse": (_I, [_P, _I]),
    "gnss_dealias": (_I, [_P]),
    "gnss_step": (_I, [_P, _I]),
    "gnss_stats": (_I, [_P, ctypes.POINTER(StatsOut), _D, _D]),
    "gnss_stats_noalloc": (_I, [_P, ctypes.POINTER(StatsOut), _D, _D]),
    "gnss_get_field": (_I, [_P, _I, _I, _D]),
    "gnss_set_field": (_I, [_P, _I, _I, _D]),
    "gnss_get_time": (_I, [_P, _D, ctypes.POINTER(ctypes.c_int64)]),
    "gnss_save_fields": (_I, [_P, ctypes.c_char_p, _I]),
    "gnss_load_fields": (_I, [_P, ctypes.c_char_p, _I]),
    "gnss_get_slice": (_I, [_P, _I, _I, _I, _D]),
    "gnss_save_slices": (_I, [_P, ctypes.c_char_p, _I]),
    "gnss_surface_area": (_I, [_P, _I, ctypes.c_double, _D]),
    "gnss_timing_enable": (_I, [_P, _I]),
    "gnss_timing_report": (_I, [_P, ctypes.c_char_p, _I]),
    "gnss_bench_fft": (_I, [_P, _I, _D, _D]),
    "gnss_bench_copy": (_I, [_P, _I, _D, _D]),
    "gnss_autotune": (_I, [_P, _I, ctypes.c_char_p, _I]),
    "gnss_set_tuning": (_I, [_P, ctypes.c_char_p]),
    "gnss_ensemble_create": (_I, [_I, ctypes.POINTER(Params), ctypes.POINTER(_P)]),
    "gnss_ensemble_destroy": (_I, [_P]),
    "gnss_ensemble_last_error": (ctypes.c_char_p, [_P]),
    "gnss_ensemble_init": (_I, [_P]),
    "gnss_ensemble_step": (_I, [_P, _I]),
    "gnss_ensemble_stats": (_I, [_P, _I, ctypes.POINTER(StatsOut), _D, _D]),
    "gnss_ensemble_get_field": (_I, [_P, _I, _I, _I, _D]),
    "gnss_ensemble_autotune": (_I, [_P, _I, ctypes.c_char_p, _I]),
}

_loaded: dict = {}


def load(path: Optional[str] = None) -> ctypes.CDLL:
    """Load a library implementing include/gnss.h (default: the in-tree MI355X engine)."""
    path = os.path.abspath(path or ENGINE_LIB)
    if path in _loaded:
        return _loaded[path]
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _loaded[path] = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


@dataclass
class Config:
    """Runtime parameters (the reference's dnsparams.h macros, dnsparams.h:9-41)."""
    n: Sequence[int] = (64, 64, 64)
    L: Sequence[float] = (2 * np.pi, 2 * np.pi, 2 * np.pi)
    re: float = 1000.0
    sc: Sequence[float] = (0.7,)
    theta: float = 0.05
    jet_h: float = 1.0
    noise_amp: float = 0.05
    noise_kmax: float = 8.0
    seed: int = 20261202
    dt: float = 1e-3
    precision: int = FP64
    flags: int = 0
    alpha: float = 0.0       # colloid drift coefficient (with FLAG_COLLOID: last scalar is the colloid)
    theta_c: float = 0.0     # colloid profile thickness (0: theta)

    def to_params(self) -> Params:
        p = Params()
        p.nx, p.ny, p.nz = (int(v) for v in self.n)
        p.lx, p.ly, p.lz = (float(v) for v in self.L)
        p.re = float(self.re)
        p.n_scalars = len(self.sc)
        for i, s in enumerate(self.sc):
            p.sc[i] = float(s)
        p.theta = float(self.theta)
        p.jet_h = float(self.jet_h)
        p.noise_amp = float(self.noise_amp)
        p.noise_kmax = float(self.noise_kmax)
        p.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        p.dt = float(self.dt)
        p.precision = int(self.precision)
        p.flags = int(self.flags)
        p.alpha = float(self.alpha)
        p.theta_c = float(self.theta_c)
        return p


class Solver:
    """One realisation: owns a gnss_ctx of the given library."""

    def __init__(self, cfg: Config, lib: Optional[ctypes.CDLL] = None):
        self.lib = lib if lib is not None else load()
        self.cfg = cfg
        self._p = cfg.to_params()
        ctx = _P()
        rc = self.lib.gnss_create(ctypes.byref(self._p), ctypes.byref(ctx))
        if rc:
            raise GnssError(rc, self.lib.gnss_last_error(None).decode())
        self.ctx = ctx
        s = Sizes()
        self._check(self.lib.gnss_get_sizes(self.ctx, ctypes.byref(s)))
        self.sizes = s
        self.nx, self.ny, self.nz = (int(v) for v in cfg.n)
        self.nzh = s.nzh
        self.S = len(cfg.sc)

    # -- lifecycle -------------------------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.gnss_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc:
            raise GnssError(rc, self.lib.gnss_last_error(self.ctx).decode())

    # -- driver entry points ---------------------------------------------------------------
    def init_wavenumbers(self):
        self._check(self.lib.gnss_init_wavenumbers(self.ctx))

    def init_jet(self):
        self._check(self.lib.gnss_init_jet(self.ctx))

    def setup(self):
        self.init_wavenumbers()
        self.init_jet()
        return self

    def step(self, n: int = 1):
        self._check(self.lib.gnss_step(self.ctx, int(n)))

    def dealias(self):
        self._check(self.lib.gnss_dealias(self.ctx))

    def forward(self, f: int):
        self._check(self.lib.gnss_forward(self.ctx, int(f)))

    def inverse(self, f: int):
        self._check(self.lib.gnss_inverse(self.ctx, int(f)))

    def stats(self, want_prof: bool = True, want_spec: bool = True, noalloc: bool = False):
        """gnss_stats (noalloc: gnss_stats_noalloc, the allocation-free form)."""
        o = StatsOut()
        yp = np.zeros((self.sizes.n_prof, self.ny)) if want_prof else None
        sp = np.zeros((1 + self.S, self.sizes.n_bins)) if want_spec else None
        fn = self.lib.gnss_stats_noalloc if noalloc else self.lib.gnss_stats
        self._check(fn(self.ctx, ctypes.byref(o), _dptr(yp) if yp is not None else None,
                       _dptr(sp) if sp is not None else None))
        return o.as_dict(self.S), yp, sp

    def save_fields(self, directory: str, it: int):
        """Checkpoint in the reference's raw physical-space format (<dir>/<u|v|w|z..>.<it>)."""
        self._check(self.lib.gnss_save_fields(self.ctx, directory.encode(), int(it)))

    def load_fields(self, directory: str, it: int):
        self._check(self.lib.gnss_load_fields(self.ctx, directory.encode(), int(it)))

    def slice(self, field: int, axis: int = 2, index: Optional[int] = None) -> np.ndarray:
        """One physical plane of `field` (FIELD_VORTICITY: |omega|); axis 2 -> [x][y] at
        z = index (default nz/2), axis 1 -> [x][z], axis 0 -> [y][z]."""
        n = self.cfg.n
        if index is None:
            index = n[axis] // 2 if axis in (0, 1, 2) else 0
        # an invalid axis gets a one-element buffer: the library rejects it before writing
        shape = {2: (n[0], n[1]), 1: (n[0], n[2]), 0: (n[1], n[2])}.get(axis, (1,))
        out = np.empty(shape)
        self._check(self.lib.gnss_get_slice(self.ctx, int(field), int(axis), int(index), _dptr(out)))
        return out

    def save_slices(self, directory: str, it: int):
        """<dir>/<name>_xy.<it>: the z = nz/2 plane of every field in the reference's format."""
        self._check(self.lib.gnss_save_slices(self.ctx, directory.encode(), int(it)))

    def surface_area(self, field: int, iso: float) -> float:
        """Area of the iso-surface field = iso (FIELD_VORTICITY: |omega| = iso)."""
        a = ctypes.c_double()
        self._check(self.lib.gnss_surface_area(self.ctx, int(field), float(iso), ctypes.byref(a)))
        return a.value

    def time(self):
        t, s = ctypes.c_double(), ctypes.c_int64()
        self._check(self.lib.gnss_get_time(self.ctx, ctypes.byref(t), ctypes.byref(s)))
        return t.value, s.value

    # -- host I/O ----------------------------------------------------------------------------
    def get_spectral(self, f: int) -> np.ndarray:
        """Spectral field, reference order [ky][kx][kz] (complex128)."""
        a = np.zeros((self.ny, self.nx, self.nzh), dtype=np.complex128)
        self._check(self.lib.gnss_get_field(self.ctx, int(f), SPECTRAL,
                                            a.ctypes.data_as(_D)))
        return a

    def set_spectral(self, f: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == (self.ny, self.nx, self.nzh), a.shape
        self._check(self.lib.gnss_set_field(self.ctx, int(f), SPECTRAL, a.ctypes.data_as(_D)))

    def get_physical(self, f: int) -> np.ndarray:
        """Physical buffer of field f, [x][y][z] (float64)."""
        a = np.zeros((self.nx, self.ny, self.nz))
        self._check(self.lib.gnss_get_field(self.ctx, int(f), PHYSICAL, _dptr(a)))
        return a

    def set_physical(self, f: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.nx, self.ny, self.nz), a.shape
        self._check(self.lib.gnss_set_field(self.ctx, int(f), PHYSICAL, _dptr(a)))

    def wavenumbers(self):
        kx, ky, kz = np.zeros(self.nx), np.zeros(self.ny), np.zeros(self.nzh)
        self._check(self.lib.gnss_get_wavenumbers(self.ctx, _dptr(kx), _dptr(ky), _dptr(kz)))
        return kx, ky, kz

    def mask(self) -> np.ndarray:
        m = np.zeros((self.ny, self.nx, self.nzh), dtype=np.uint8)
        self._check(self.lib.gnss_get_mask(self.ctx, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return m

    # -- engine-only instrumentation ---------------------------------------------------------
    def timing(self, enable: bool):
        self._check(self.lib.gnss_timing_enable(self.ctx, 1 if enable else 0))

    def timing_report(self) -> dict:
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self.lib.gnss_timing_report(self.ctx, buf, len(buf)))
        return json.loads(buf.value.decode() or "{}")

    def autotune(self, reps: int = 5) -> dict:
        """Time the compiled pass variants on this context and keep the fastest."""
        buf = ctypes.create_string_buffer(1 << 14)
        self._check(self.lib.gnss_autotune(self.ctx, int(reps), buf, len(buf)))
        return json.loads(buf.value.decode())

    def set_tuning(self, spec: str):
        self._check(self.lib.gnss_set_tuning(self.ctx, spec.encode()))

    def bench_copy(self, reps: int = 10):
        """Mean ms of a plain device copy of one spectral field, and the bytes it reads + writes."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.gnss_bench_copy(self.ctx, int(reps), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def bench_fft(self, reps: int = 10):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.gnss_bench_fft(self.ctx, int(reps), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Ensemble:
    """Batched realisations of uneven sizes advanced together (gnss_ensemble_*)."""

    def __init__(self, cfgs: Sequence[Config], lib: Optional[ctypes.CDLL] = None):
        self.lib = lib if lib is not None else load()
        self.cfgs = list(cfgs)
        arr = (Params * len(self.cfgs))(*[c.to_params() for c in self.cfgs])
        self._arr = arr
        e = _P()
        rc = self.lib.gnss_ensemble_create(len(self.cfgs), arr, ctypes.byref(e))
        if rc:
            raise GnssError(rc, self.lib.gnss_ensemble_last_error(None).decode())
        self.ens = e

    def _check(self, rc):
        if rc:
            raise GnssError(rc, self.lib.gnss_ensemble_last_error(self.ens).decode())

    def close(self):
        if getattr(self, "ens", None):
            self.lib.gnss_ensemble_destroy(self.ens)
            self.ens = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        self._check(self.lib.gnss_ensemble_init(self.ens))

    def step(self, n: int = 1):
        self._check(self.lib.gnss_ensemble_step(self.ens, int(n)))

    def stats(self, i: int):
        c = self.cfgs[i]
        S = len(c.sc)
        o = StatsOut()
        self._check(self.lib.gnss_ensemble_stats(self.ens, int(i), ctypes.byref(o), None, None))
        return o.as_dict(S)

    def stats_full(self, i: int):
        c = self.cfgs[i]
        S = len(c.sc)
        o = StatsOut()
        self._check(self.lib.gnss_ensemble_stats(self.ens, int(i), ctypes.byref(o), None, None))
        return o.as_dict(S)

    def get_spectral(self, i: int, f: int) -> np.ndarray:
        c = self.cfgs[i]
        nx, ny, nz = c.n
        a = np.zeros((ny, nx, nz // 2 + 1), dtype=np.complex128)
        self._check(self.lib.gnss_ensemble_get_field(self.ens, int(i), int(f), SPECTRAL,
                                                     a.ctypes.data_as(_D)))
        return a

    def get_physical(self, i: int, f: int) -> np.ndarray:
        """Member i's field f in physical space, [x][y][z], from its current spectral state."""
        nx, ny, nz = self.cfgs[i].n
        a = np.zeros((nx, ny, nz), dtype=np.float64)
        self._check(self.lib.gnss_ensemble_get_field(self.ens, int(i), int(f), PHYSICAL,
                                                     a.ctypes.data_as(_D)))
        return a

    def autotune(self, reps: int = 5) -> list:
        """Tunes the batched engines (their "<key>b" table entries); returns one report per
        batched engine."""
        buf = ctypes.create_string_buffer(1 << 20)
        self._check(self.lib.gnss_ensemble_autotune(self.ens, int(reps), buf, len(buf)))
        return json.loads(buf.value.decode())


# The five workloads the engine is measured on (synthetic seeded jets, L = 2 pi).
def config(name: str, **over) -> Config:
    two_pi = 2 * np.pi
    base = dict(L=(two_pi,) * 3, jet_h=1.0, noise_amp=0.05, seed=20261202)
    table = {
        "jet64": dict(n=(64,) * 3, theta=0.05, re=1000.0, sc=(0.7,), noise_kmax=8.0, dt=1e-3),
        "jet256": dict(n=(256,) * 3, theta=0.03, re=3000.0, sc=(0.7, 2.0), noise_kmax=16.0,
                       dt=5e-4),
        "jet512": dict(n=(512,) * 3, theta=0.02, re=5000.0, sc=(1.0,), noise_kmax=32.0,
                       dt=2.5e-4),
        "jet1024": dict(n=(1024,) * 3, theta=0.01, re=10000.0, sc=(1.0,), noise_kmax=64.0,
                        dt=1e-4),
    }
    d = dict(base)
    d.update(table[name])
    d.update(over)
    return Config(**d)


def with_colloid(cfg: Config, alpha: float = 0.0033, sc_c: float = 0.7, theta_c: float = 0.0) -> Config:
    """`cfg` plus a colloid as the last scalar, carried by u + alpha grad(scalar 0); defaults are
    the reference's alpha, Sc_c and theta_c = theta (dnsparams.h:28-33)."""
    d = dict(cfg.__dict__)
    d["sc"] = tuple(cfg.sc) + (sc_c,)
    d["flags"] = int(cfg.flags) | FLAG_COLLOID
    d["alpha"] = float(alpha)
    d["theta_c"] = float(theta_c)
    return Config(**d)


def _u01(seed: int, k: int) -> float:
    """Portable uniform draw in [0, 1) from a SplitMix64 hash of (seed, k)."""
    m = 0xFFFFFFFFFFFFFFFF
    z = (seed * 0x100000001B3 + k + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return (z >> 11) / float(1 << 53)


def ensemble_configs(seed: int = 20261202, counts=((64, 32), (96, 16), (128, 8)),
                     precision: int = FP32, dt: float = 1e-3) -> list:
    """Uneven batched ensemble: `count` realisations per grid size n, member i with seed + i,
    theta ~ U[0.03, 0.08] and Re ~ U[500, 2000] drawn from that seed, one scalar Sc = 0.7,
    noise rms 0.05 on |m| <= n/8."""
    out = []
    i = 0
    for n, count in counts:
        for _ in range(count):
            s = seed + i
            out.append(Config(n=(n, n, n), re=500.0 + 1500.0 * _u01(s, 2), sc=(0.7,),
                              theta=0.03 + 0.05 * _u01(s, 1), noise_amp=0.05, noise_kmax=n / 8.0,
                              seed=s, dt=dt, precision=precision))
            i += 1
    return out
