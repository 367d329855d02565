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
