"""CPU checks of the host-side driver: particle initialisation and the binary
particle file (m_particles.f90:157-337, 558-744) and the scalar statistics
text formats (m_stats.f90:62-80, 421-430).  No GPU call is made."""
import math
import os

import numpy as np

from hit3d_amd import driver


def test_particle_lattice_matches_oracle(oracle):
    for nptot, n in ((1100, 32), (64, 16), (1_000_000, 64)):
        ip, xyz = driver.particles_init(nptot, n)
        want_n, want = oracle.particles_lattice(nptot, n)
        assert len(ip) == want_n and np.array_equal(ip, np.arange(1, want_n + 1))
        assert np.array_equal(xyz, want)
        assert (xyz >= 1.0).all() and (xyz < n + 1.0).all()


def test_particles_from_text_file_wrap(tmp_path):
    """<run>.pt positions outside [0, 2 pi) are wrapped (m_particles.f90:283-290)."""
    n = 16
    two_pi = 2 * math.pi
    pts = np.array([[0.1, 1.0, 2.0], [two_pi + 0.5, -0.25, 3.0], [6.0, 6.2, -two_pi - 1.0]])
    path = tmp_path / "abcdefghij.pt"
    path.write_text("\n".join(" ".join(f"{v:.17g}" for v in row) for row in pts) + "\n")
    ip, xyz = driver.particles_init(3, n, str(path))
    wrapped = np.where((pts >= two_pi) | (pts < 0), pts - two_pi * np.floor(pts / two_pi), pts)
    assert np.array_equal(xyz, wrapped / (two_pi / n) + 1.0)
    assert list(ip) == [1, 2, 3]


def test_particle_binary_layout_and_round_trip(tmp_path):
    """int32 nptot, then (int32 id, 3 x FP64 (x-1)*dx) per particle, then
    FP64 time, with no record markers (stream access)."""
    n = 32
    rng = np.random.default_rng(3)
    xyz = rng.uniform(1.0, n + 1.0, (57, 3))
    ip = rng.permutation(np.arange(1, 58)).astype(np.int32)
    path = tmp_path / "pt.000010"
    driver.write_particles_binary(str(path), ip, xyz, n, 1.25)
    raw = path.read_bytes()
    assert len(raw) == 4 + 57 * 28 + 8
    assert np.frombuffer(raw[:4], "<i4")[0] == 57
    assert np.frombuffer(raw[4:8], "<i4")[0] == ip[0]
    dx = 2 * math.pi / n
    assert np.array_equal(np.frombuffer(raw[8:32], "<f8"), (xyz[0] - 1.0) * dx)
    assert np.frombuffer(raw[-8:], "<f8")[0] == 1.25
    ip2, xyz2, t = driver.read_particles_binary(str(path), n)
    assert np.array_equal(ip2, ip) and t == 1.25
    assert np.abs(xyz2 - xyz).max() < 1e-12


def test_scalar_stats_text_formats(tmp_path):
    st = {"sc_diss": np.array([0.5, 2.0]), "mean": np.array([0.0, -1.0]),
          "variance": np.array([1.0, 3.0]), "sc_min": np.array([-2.0, -4.0]),
          "sc_max": np.array([2.0, 4.0]), "es_spec": np.array([[1.0, 0.5], [0.25, 0.125]])}
    driver.write_scalar_stats(st, 7, 0.5, str(tmp_path))
    driver.write_scalar_stats(st, 8, 0.6, str(tmp_path))
    lines = (tmp_path / "sc02.gp").read_text().splitlines()
    assert lines[0].startswith("# 1.itime 2.time")
    assert len(lines) == 3
    assert lines[1] == ("      7   0.500000E+00   0.200000E+01  -0.100000E+01   0.300000E+01"
                        "  -0.400000E+01   0.400000E+01")
    es = (tmp_path / "es_01.gp").read_text().splitlines()
    assert es[:3] == ["", "", "# ITIME=      7 TIME=   0.50000000E+00"]
    assert es[3] == "   1   0.100000E+01"


def test_fortran_e_editing():
    """Ew.d as gfortran writes it (mantissa in [0.1, 1), rounded to nearest)."""
    f = driver.fortran_e
    assert f(0.0) == "   0.000000E+00"
    assert f(1234.5678) == "   0.123457E+04"
    assert f(-0.000123456789) == "  -0.123457E-03"
    assert f(9.9999996) == "   0.100000E+02"
    assert f(1e-120) == "   0.100000-119"
    assert f(0.5, 17, 8) == "   0.50000000E+00"


def test_restart_file_layout_and_round_trip(tmp_path):
    """36-byte header (int32 nx, ny, nz, nums, nles; FP64 time, dt), then
    each field's n^3 spectral values with the Nyquist pair dropped."""
    n = 8
    rng = np.random.default_rng(11)
    f = rng.standard_normal((4, n, n, n + 2))
    path = tmp_path / "abcdefghij.64.000020"
    driver.write_restart(str(path), f, n, 1, 0, 0.75, 0.002)
    raw = path.read_bytes()
    assert len(raw) == 36 + 4 * n ** 3 * 8
    assert list(np.frombuffer(raw[:20], "<i4")) == [n, n, n, 1, 0]
    assert list(np.frombuffer(raw[20:36], "<f8")) == [0.75, 0.002]
    assert np.array_equal(np.frombuffer(raw[36:36 + 8 * n], "<f8"), f[0, 0, 0, :n])
    hdr, g = driver.read_restart(str(path), n)
    assert hdr["nums"] == 1 and hdr["time"] == 0.75 and hdr["dt"] == 0.002
    assert np.array_equal(g[:, :, :, :n], f[:, :, :, :n])
    assert not g[:, :, :, n:].any()


def test_tmp4_layout(tmp_path):
    a = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)   # [z, y, x]
    path = tmp_path / "u.000000"
    driver.write_tmp4(str(path), a)
    raw = path.read_bytes()
    assert list(np.frombuffer(raw[:12], "<i4")) == [4, 3, 2]
    assert np.array_equal(np.frombuffer(raw[12:], "<f4"), a.ravel())


def test_closure_file_round_trip(tmp_path):
    """<run_name>.sgs: parameters, scaling and the blend weight come back
    unchanged; a file without the weight (32 doubles) reads as weight 0."""
    from hit3d_amd.driver import read_closure_file, write_closure_file
    rng = np.random.default_rng(1)
    params = rng.standard_normal(6288).astype(np.float32)
    norm = (rng.standard_normal(10), 1 + rng.random(10), rng.standard_normal(6), 1 + rng.random(6))
    path = tmp_path / "a.sgs"
    write_closure_file(str(path), params, norm, mix=0.5)
    assert path.stat().st_size == 6288 * 4 + 33 * 8
    p2, n2, mix = read_closure_file(str(path))
    assert np.array_equal(p2, params) and mix == 0.5
    assert all(np.array_equal(a, b) for a, b in zip(n2, norm))
    old = tmp_path / "b.sgs"
    old.write_bytes(path.read_bytes()[:-8])
    assert read_closure_file(str(old))[2] == 0.0
