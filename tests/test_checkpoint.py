"""Checkpoint / restart in the reference's raw format (iofuncs.cu:329-358 writer, 462-497 reader):
an 8-byte header holding the byte count as a double, then nx*ny*nz doubles in [x][y][z] order,
one file per variable (u, v, w, z) named <var>.<iter>."""
import os

import numpy as np
import pytest


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def read_reference_file(path, n):
    raw = np.fromfile(path, dtype=np.float64)
    assert raw[0] == n[0] * n[1] * n[2] * 8  # header: byte count as a double
    return raw[1:].reshape(n)


CFG = dict(n=(16, 24, 32), re=300.0, sc=(0.7,), theta=0.1, noise_kmax=5.0, noise_amp=0.2, dt=2e-3)


def test_oracle_writes_reference_format(g, oracle, tmp_path):
    s = g.Solver(g.Config(**CFG), oracle).setup()
    s.step(3)
    s.save_fields(str(tmp_path), 3)
    assert sorted(os.listdir(tmp_path)) == ["gnss_state.3", "u.3", "v.3", "w.3", "z.3"]
    n = CFG["n"]
    for f, name in enumerate("uvwz"):
        x = read_reference_file(tmp_path / f"{name}.3", n)
        spec = s.get_spectral(f).transpose(1, 0, 2)  # -> [kx][ky][kz]
        assert np.abs(x - np.fft.irfftn(spec, s=n, axes=(0, 1, 2))).max() < 1e-14
    assert open(tmp_path / "gnss_state.3").read().split()[:2] == ["step", "3"]


def test_oracle_restart_continues_the_run(g, oracle, tmp_path):
    a = g.Solver(g.Config(**CFG), oracle).setup()
    a.step(4)
    a.save_fields(str(tmp_path), 4)
    a.step(4)
    b = g.Solver(g.Config(**CFG), oracle)
    b.init_wavenumbers()
    b.load_fields(str(tmp_path), 4)
    assert b.time()[1] == 4
    b.step(4)
    for f in range(4):
        assert rel(b.get_spectral(f), a.get_spectral(f)) < 1e-13


def test_size_header_is_checked(g, oracle, tmp_path):
    s = g.Solver(g.Config(**CFG), oracle).setup()
    s.save_fields(str(tmp_path), 0)
    raw = np.fromfile(tmp_path / "v.0", dtype=np.float64)
    raw[0] = 8.0 * 10
    raw.tofile(tmp_path / "v.0")
    t = g.Solver(g.Config(**CFG), oracle)
    t.init_wavenumbers()
    with pytest.raises(g.GnssError) as e:
        t.load_fields(str(tmp_path), 0)
    assert e.value.code == 1 and "size header" in str(e.value)


@pytest.mark.gpu
def test_engine_and_oracle_read_each_others_files(g, engine, oracle, tmp_path):
    a = g.Solver(g.Config(**CFG), engine).setup()
    a.step(3)
    a.save_fields(str(tmp_path), 3)
    o = g.Solver(g.Config(**CFG), oracle)
    o.init_wavenumbers()
    o.load_fields(str(tmp_path), 3)
    for f in range(4):
        assert rel(o.get_spectral(f), a.get_spectral(f)) < 1e-13
    o.step(2)
    o.save_fields(str(tmp_path), 5)
    b = g.Solver(g.Config(**CFG), engine)
    b.init_wavenumbers()
    b.load_fields(str(tmp_path), 5)
    assert b.time()[1] == 5
    for f in range(4):
        assert rel(b.get_spectral(f), o.get_spectral(f)) < 1e-13
    # restart continues the engine run
    a.step(2)
    a.step(3)
    b.step(3)
    for f in range(4):
        assert rel(b.get_spectral(f), a.get_spectral(f)) < 1e-12
