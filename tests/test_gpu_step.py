"""GPU checks of the fused time step (hit3d_step): against the unfused
sequence of reference subroutines, against the oracle, and the error of the
mixed FP32/FP64 form.

Tolerances: fused vs unfused FP64 1e-13 relative per step; mixed vs FP64
single-step RHS 1e-5 relative (the stated bound for the mixed form).
"""
import sys

import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS, RHS_OLD

pytestmark = pytest.mark.gpu

RN1 = 67748937.0


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def note(msg):
    print(msg, flush=True)


def fresh(n, forced=True, nu=1 / 800.0):
    h = hit3d_amd.Hit3D(n, dealias=0)
    h.set_physics(nu, flow_type=1 if forced else 0, kfmax=2, famp=0.5)
    h.init_velocity(RN1, 2, 4, 3.0)
    h.dealias_all()
    return h


def fields(h, arr=FIELDS):
    return np.stack([h.get(arr, c) for c in (1, 2, 3)])


@pytest.mark.parametrize("n", [16, 32, 64, 128])
def test_fused_matches_unfused(n):
    note(f"fused vs unfused n={n}")
    h = fresh(n)
    f0 = fields(h)
    dt = 0.002
    res = {}
    for fused in (True, False):
        for c in range(3):
            h.put(FIELDS, c + 1, f0[c])
        h.set_fusion(fused)
        fov = True
        out = []
        for _ in range(3):
            cr, en, fov, _ = h.step(dt, fov, True)
            out.append((cr, en, fields(h), fields(h, RHS_OLD)))
        res[fused] = out
    for (c1, e1, f1, r1), (c2, e2, f2, r2) in zip(res[True], res[False]):
        assert abs(c1 - c2) <= 1e-13 * c2
        assert abs(e1 - e2) <= 1e-13 * e2
        assert rel(f1, f2) < 1e-13
        assert rel(r1, r2) < 1e-13
    h.close()


def test_fused_forced_trajectory_vs_oracle(oracle):
    n = 32
    note("fused forced trajectory n=32")
    cfg = oracle.Config(n=n, nu=1 / 400.0, dt=0.001, dealias=0, flow_type=1, kfmax=2, famp=0.5,
                        variable_dt=True)
    h = fresh(n, nu=cfg.nu)
    f0 = fields(h)
    st = oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt)
    from hit3d_amd.driver import my_dt
    fov = True
    dt = cfg.dt
    for it in range(8):
        c, e, fov, _ = h.step(dt, fov, True)
        st = oracle.step(cfg, st)
        assert abs(c - st.courant) < 1e-12 * st.courant, it
        assert abs(e - st.energy) < 1e-12 * st.energy, it
        dt = my_dt(dt, c)
        assert dt == st.dt
    assert rel(fields(h), st.fields[:3]) < 1e-11
    h.close()


@pytest.mark.parametrize("n", [32, 128])
def test_mixed_single_rhs_error(n):
    """The mixed form's first-step RHS (left in rhs_old by the Euler step)
    agrees with the FP64 one to 1e-5 relative."""
    note(f"mixed rhs n={n}")
    h = fresh(n)
    f0 = fields(h)
    rhs = {}
    for prec in (0, 1):
        for c in range(3):
            h.put(FIELDS, c + 1, f0[c])
        h.step(0.001, True, True, precision=prec)
        rhs[prec] = fields(h, RHS_OLD)
    err = rel(rhs[1], rhs[0])
    note(f"mixed single-RHS rel err {err:.3e}")
    assert err < 1e-5
    assert err > 1e-12  # really computed in FP32
    h.close()


def test_mixed_short_run_spectrum():
    """After 20 mixed steps E(k) stays within 1e-4 of the FP64 run."""
    n = 64
    note("mixed spectrum n=64")
    h = fresh(n)
    f0 = fields(h)
    spec = {}
    for prec in (0, 1):
        for c in range(3):
            h.put(FIELDS, c + 1, f0[c])
        fov = True
        for _ in range(20):
            _, _, fov, _ = h.step(0.002, fov, True, precision=prec)
        spec[prec] = h.stats()[0]
    sig = spec[0] > 1e-10 * spec[0].max()
    assert np.max(np.abs(spec[1][sig] - spec[0][sig]) / spec[0][sig]) < 1e-4
    h.close()


def test_step_rejects_bad_precision():
    h = hit3d_amd.Hit3D(16)
    h.set_physics(0.01)
    with pytest.raises(hit3d_amd.Hit3dError, match="precision"):
        h.step(0.001, True, True, precision=7)
    h.close()


def test_second_open_instance_refused():
    """One grid per process: a second Hit3D while one is open is refused
    (it would otherwise re-initialise the state the first one drives), and
    allowed once the first is closed."""
    h = hit3d_amd.Hit3D(16)
    try:
        with pytest.raises(hit3d_amd.Hit3dError, match="one grid per process"):
            hit3d_amd.Hit3D(32)
        assert h.n == 16
        h.set_physics(0.01)
        h.step(0.001, True, True)
    finally:
        h.close()
    h2 = hit3d_amd.Hit3D(32)
    h2.close()


def test_mixed_rhs_old_storage_and_transitions():
    """The mixed step keeps rhs_old(1:3) in FP32 between steps.  Read back
    after a mixed step it holds FP32-representable values and zeros outside
    the 2/3 cube; after an FP64 step, FP64 values.  Reading it between steps
    (a conversion to FP64 and, at the next mixed step, back) must leave a
    trajectory that switches precisions bit-identical to the same trajectory
    without the reads."""
    n = 32
    note("mixed rhs_old storage and precision switches n=32")
    h = fresh(n)
    f0 = fields(h)
    r0 = fields(h, RHS_OLD)
    seq = [1, 1, 0, 1, 1, 0, 0, 1]

    def run(read_between):
        for c in range(3):
            h.put(FIELDS, c + 1, f0[c])
            h.put(RHS_OLD, c + 1, r0[c])
        fov = True
        seen = []
        for i, p in enumerate(seq):
            _, _, fov, _ = h.step(0.002, fov, True, precision=p)
            if read_between:
                seen.append((p, fields(h, RHS_OLD)))
        return fields(h), fields(h, RHS_OLD), seen

    fa, ra, _ = run(False)
    fb, rb, seen = run(True)
    assert np.array_equal(fa, fb) and np.array_equal(ra, rb)
    K = n // 3
    j = np.arange(n)
    kept_j = (j <= K) | (j >= n - K)
    kept_x = np.repeat(np.arange(n // 2 + 1) <= K, 2)
    cube = kept_j[:, None, None] & kept_j[None, :, None] & kept_x[None, None, :]
    for p, ro in seen:
        as32 = ro.astype(np.float32).astype(np.float64)
        if p == 1:
            assert np.array_equal(ro, as32)
            assert not np.any(ro[:, ~cube])
        else:
            assert not np.array_equal(ro, as32)
    h.close()
