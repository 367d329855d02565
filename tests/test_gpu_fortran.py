"""The Fortran drop-in: hit3d_gpu (built with amdflang, -i8 -r8) reads a
reference-format input file, runs the driver loop on the GPU through the
ISO_C_BINDING module and writes the reference's stat1.gp / es.gp and restart
layout.  Its statistics must equal the Python driver's on the same library."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hit3d_amd", "fortran", "hit3d_gpu")

INPUT = """======================================================================
INPUT FILE (decaying 32^3, fixed DT, 2/3 rule)
======================================================================
32 32 32        dimensions of the grid
----------------------------------------------------------------------
000000          ITMIN
12              ITMAX
4               IPRNT1
1000            IPRNT2
1000            IWRITE4
----------------------------------------------------------------------
1000000.0       TMAX
1000000.0 0     TRESCALE, NRESCALE
0.0             TSCALAR
----------------------------------------------------------------------
0               flow type
----------------------------------------------------------------------
100.0           RE
-0.01           DT
----------------------------------------------------------------------
2               ISPCV1
4               mv1
3.0             wm0v1
----------------------------------------------------------------------
1               force_type
2               KFMAX
0.5             FAMP
----------------------------------------------------------------------
0               dealias
----------------------------------------------------------------------
0               det_rand
67748937.       RN1
93712139.       RN2
10765397.       RN3
----------------------------------------------------------------------
0               np
1               particle_tracking
0.1             time_p
0.3             particle_filter_size
----------------------------------------------------------------------
0               les_model
----------------------------------------------------------------------
0               nums
----------------------------------------------------------------------
"""


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_matches_python_driver(tmp_path):
    run = "c1decay032"
    (tmp_path / f"{run}.in").write_text(INPUT)
    print("fortran driver 32^3", flush=True)
    r = subprocess.run([EXE, run], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    stat1 = np.loadtxt(tmp_path / "stat1.gp", comments="#")
    assert stat1.shape == (3, 7)
    assert list(stat1[:, 0]) == [4, 8, 12]
    # the Python driver on the same library
    from hit3d_amd.driver import Run, read_input_file
    p = read_input_file(str(tmp_path / f"{run}.in"))
    assert p.variable_dt is False and p.dt == 0.01
    d = Run(p)
    d.begin_new()
    energies = []
    for it in range(1, 13):
        d.step()
        if it % 4 == 0:
            energies.append(d.stats().energy)
    d.h.close()
    # stat1.gp carries 6 significant digits (m_stats.f90:160)
    assert np.allclose(stat1[:, 2], energies, rtol=1e-5, atol=0)
    # restart file: 36-byte header + 3 fields of 32^3 doubles
    rs = tmp_path / f"{run}.64.000012"
    raw = rs.read_bytes()
    assert len(raw) == 36 + 3 * 32 ** 3 * 8
    hdr = np.frombuffer(raw[:20], np.int32)
    assert list(hdr) == [32, 32, 32, 0, 0]
    t, dt = np.frombuffer(raw[20:36], np.float64)
    assert abs(t - 0.12) < 1e-12 and dt == 0.01


def _input(n, itmin, itmax, iprint1, iprint2, iwrite4, flow_type, re, dt, dealias, nptot=0,
           ptrack=1, les_model=0, scalars=()):
    lines = ["=" * 70, "INPUT FILE", "=" * 70, f"{n} {n} {n}        grid", "-" * 70,
             f"{itmin:06d}          ITMIN", f"{itmax}              ITMAX",
             f"{iprint1}               IPRNT1", f"{iprint2}               IPRNT2",
             f"{iwrite4}               IWRITE4", "-" * 70, "1000000.0       TMAX",
             "1000000.0 0     TRESCALE, NRESCALE", "0.0             TSCALAR", "-" * 70,
             f"{flow_type}               flow type", "-" * 70, f"{re}           RE",
             f"{dt}           DT", "-" * 70, "2               ISPCV1", "4               mv1",
             "3.0             wm0v1", "-" * 70, "1               force_type",
             "2               KFMAX", "0.5             FAMP", "-" * 70,
             f"{dealias}               dealias", "-" * 70, "0               det_rand",
             "67748937.       RN1", "93712139.       RN2", "10765397.       RN3", "-" * 70,
             f"{nptot}               np", f"{ptrack}               particle_tracking",
             "0.0             time_p", "0.0             particle_filter_size", "-" * 70,
             f"{les_model}               les_model", "-" * 70,
             f"{len(scalars)}               nums", "-" * 70]
    if scalars:
        lines.append("type sc ir_exp peak reac")
        lines += [f"{t} {sc} 0 1.0 0.0" for t, sc in scalars]
    return "\n".join(lines) + "\n"


def _run_both(tmp_path, run, text, closure=None, seed_files=()):
    """The Fortran driver in tmp/f and the Python driver's Run.run in tmp/p
    on the same input (and the same starting files)."""
    from hit3d_amd.driver import Run, read_input_file
    out = {}
    for side in ("f", "p"):
        d = tmp_path / side
        d.mkdir(parents=True, exist_ok=True)
        (d / f"{run}.in").write_text(text)
        for name, data in seed_files:
            (d / name).write_bytes(data)
        if closure is not None:
            (d / f"{run}.sgs").write_bytes(closure)
        if side == "f":
            r = subprocess.run([EXE, run], cwd=d, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stdout + r.stderr
        else:
            p = read_input_file(str(d / f"{run}.in"))
            rr = Run(p, closure=str(d / f"{run}.sgs") if closure is not None else None)
            try:
                rr.run(run, str(d))
            finally:
                rr.h.close()
        out[side] = d
    return out["f"], out["p"]


def _same_files(fdir, pdir, names):
    for name in names:
        a, b = (fdir / name).read_bytes(), (pdir / name).read_bytes()
        assert a == b, name


def _same_stats(fdir, pdir, name="stat1.gp"):
    a = np.loadtxt(fdir / name, comments="#")
    b = np.loadtxt(pdir / name, comments="#")
    assert a.shape == b.shape and np.array_equal(a[:, 0], b[:, 0])
    assert np.allclose(a[:, 1:], b[:, 1:], rtol=2e-6, atol=0)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_phase_shift_scalars_and_restart(tmp_path):
    """A 16^3 miniature of the 256^3 configuration (forced, phase shift, two
    type-0 scalars, variable DT) through the Fortran driver and the Python
    driver: restart files at IPRINT2 and at the end, snapshots (velocity
    and scalars) bit for bit, statistics to their printed digits.  Then a
    restart from ITMIN = 4 (dt halved, begin_restart.f90:33-38) through
    both drivers, again bit for bit."""
    print("fortran driver: phase shift + scalars, restart", flush=True)
    run = "c2mini0016"
    text = _input(16, 0, 8, 2, 4, 4, 1, 400.0, 0.001, 1, scalars=((0, 0.7), (0, 1.0)))
    f, p = _run_both(tmp_path / "a", run, text)
    _same_files(f, p, [f"{run}.64.000004", f"{run}.64.000008", "u.000000", "w.000004",
                       "sc01.000008", "sc02.000004"])
    _same_stats(f, p)
    _same_stats(f, p, "sc01.gp")
    raw = (f / f"{run}.64.000004").read_bytes()
    assert list(np.frombuffer(raw[:20], np.int32)) == [16, 16, 16, 2, 0]
    text2 = _input(16, 4, 8, 2, 4, 4, 1, 400.0, 0.001, 1, scalars=((0, 0.7), (0, 1.0)))
    f2, p2 = _run_both(tmp_path / "b", run, text2, seed_files=[(f"{run}.64.000004", raw)])
    _same_files(f2, p2, [f"{run}.64.000008", "u.000008"])
    dt_file = np.frombuffer(raw[28:36], np.float64)[0]
    # the restarted run starts from dt/2, so it differs from the uninterrupted one
    assert (f2 / f"{run}.64.000008").read_bytes() != (f / f"{run}.64.000008").read_bytes()
    assert dt_file > 0.0


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_les_particles_snapshots(tmp_path):
    """A 32^3 Smagorinsky LES (les_model 1) with 4096 CINT particles on the
    lattice: restart, velocity and nu_t snapshots and the particle files
    pt.NNNNNN from the Fortran driver equal the Python driver's bit for
    bit."""
    print("fortran driver: LES + particles", flush=True)
    run = "c4mini0032"
    text = _input(32, 0, 6, 3, 1000, 3, 1, 800.0, -0.002, 0, nptot=4096, ptrack=1,
                  les_model=1)
    f, p = _run_both(tmp_path, run, text)
    _same_files(f, p, [f"{run}.64.000006", "u.000003", "nu_t.000003", "nu_t.000006",
                       "pt.000003", "pt.000006"])
    _same_stats(f, p)
    raw = (f / "pt.000006").read_bytes()
    assert np.frombuffer(raw[:4], np.int32)[0] == 4096 and len(raw) == 4 + 4096 * 28 + 8


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_learned_closure(tmp_path):
    """les_model 11 with the closure read from <run_name>.sgs (its stress
    blended half and half with Smagorinsky's): the Fortran and Python
    drivers give the same restart file."""
    from hit3d_amd.driver import write_closure_file
    print("fortran driver: learned closure", flush=True)
    rng = np.random.default_rng(3)
    params = (rng.standard_normal(6288) * 0.05).astype(np.float32)
    norm = (np.zeros(10), np.ones(10), np.zeros(6), np.full(6, 1e-3))
    path = tmp_path / "closure.sgs"
    write_closure_file(str(path), params, norm, mix=0.5)
    run = "c4mlp00032"
    text = _input(32, 0, 4, 2, 1000, 1000, 1, 800.0, -0.002, 0, les_model=11)
    f, p = _run_both(tmp_path, run, text, closure=path.read_bytes())
    _same_files(f, p, [f"{run}.64.000004"])
    _same_stats(f, p)
