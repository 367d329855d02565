"""Two options of the fused 2/3-rule step for the nonlinear term, through
the C ABI against the oracle's variants: the rotational form u x omega
(oracle.rhs_velocity_rotational) and the divergence of the trace-free
product tensor (oracle.rhs_velocity_deviatoric), both pinned by closed
forms and by the identity that their projection equals that of the
divergence form for band-limited fields.

* The first step's rhs_old holds the unprojected RHS, which tells the forms
  apart: equal to the oracle's variant plus forcing (1e-12), and different
  from the divergence form's by a gradient.
* A forced trajectory equals the oracle's trajectory of the same form
  (1e-11), and the divergence form's from the same field (the projection
  removes the gradient part): fields within 1e-11 in FP64; each mixed form
  within 1e-5 of its FP64 step.
"""
import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS, RHS_OLD

pytestmark = pytest.mark.gpu

RN1 = 67748937.0


def note(msg):
    print(msg, flush=True)


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def seeded(h):
    h.init_velocity(RN1, 2, 4, 3.0)
    h.dealias_all()
    return np.stack([h.get(FIELDS, c) for c in (1, 2, 3)])


@pytest.mark.parametrize("n", [32, 64])
def test_rotational_rhs_and_trajectory_match_oracle(oracle, n):
    note(f"rotational form n={n}")
    cfg = oracle.Config(n=n, nu=1 / 400.0, dt=0.002, dealias=0, flow_type=1, kfmax=2,
                        famp=0.5, rotational=True)
    h = hit3d_amd.Hit3D(n)
    try:
        h.set_physics(cfg.nu, flow_type=1, kfmax=2, famp=0.5)
        h.set_nonlinear(1)
        f0 = seeded(h)
        # first step: rhs_old = the unprojected rotational RHS + forcing
        c, e, fov, _ = h.step(cfg.dt, True, True)
        r, _ = oracle.rhs_velocity_rotational(f0, cfg.nu, n)
        r, e_ref = oracle.force_velocity(f0, r, 2, 0.5, n)
        got = np.stack([h.get(RHS_OLD, k) for k in (1, 2, 3)])
        assert rel(got, r) < 1e-12
        assert abs(e - e_ref) < 1e-12 * e_ref
        rd, _ = oracle.rhs_velocity(f0, cfg.nu, 0, n)
        assert rel(got, oracle.force_velocity(f0, rd, 2, 0.5, n)[0]) > 1e-3
        st = oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt)
        st = oracle.step(cfg, st)
        assert abs(c - st.courant) < 1e-12 * st.courant
        for it in range(4):
            c, e, fov, _ = h.step(cfg.dt, fov, True)
            st = oracle.step(cfg, st)
            assert abs(c - st.courant) < 1e-12 * st.courant, it
            assert abs(e - st.energy) < 1e-12 * st.energy, it
        got = np.stack([h.get(FIELDS, k) for k in (1, 2, 3)])
        assert rel(got, st.fields) < 1e-11
    finally:
        h.close()


@pytest.mark.parametrize("n", [32, 64])
def test_deviatoric_rhs_and_trajectory_match_oracle(oracle, n):
    """Form 2 (the trace-free tensor): the first step's rhs_old equals the
    oracle's deviatoric RHS plus forcing (1e-12), differs from the
    divergence form's, and a forced trajectory equals the oracle's."""
    note(f"deviatoric form n={n}")
    cfg = oracle.Config(n=n, nu=1 / 400.0, dt=0.002, dealias=0, flow_type=1, kfmax=2,
                        famp=0.5, deviatoric=True)
    h = hit3d_amd.Hit3D(n)
    try:
        h.set_physics(cfg.nu, flow_type=1, kfmax=2, famp=0.5)
        h.set_nonlinear(2)
        f0 = seeded(h)
        c, e, fov, _ = h.step(cfg.dt, True, True)
        r, _ = oracle.rhs_velocity_deviatoric(f0, cfg.nu, n)
        r, e_ref = oracle.force_velocity(f0, r, 2, 0.5, n)
        got = np.stack([h.get(RHS_OLD, k) for k in (1, 2, 3)])
        assert rel(got, r) < 1e-12
        rd, _ = oracle.rhs_velocity(f0, cfg.nu, 0, n)
        assert rel(got, oracle.force_velocity(f0, rd, 2, 0.5, n)[0]) > 1e-3
        st = oracle.step(cfg, oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt))
        assert abs(c - st.courant) < 1e-12 * st.courant
        for it in range(4):
            c, e, fov, _ = h.step(cfg.dt, fov, True)
            st = oracle.step(cfg, st)
            assert abs(c - st.courant) < 1e-12 * st.courant, it
            assert abs(e - st.energy) < 1e-12 * st.energy, it
        got = np.stack([h.get(FIELDS, k) for k in (1, 2, 3)])
        assert rel(got, st.fields) < 1e-11
    finally:
        h.close()


def test_rotational_equals_divergence_and_mixed(oracle):
    """128^3: five fused steps of each form from one field (graph replay on)
    give the same fields to 1e-11; the mixed rotational and deviatoric
    forms are within 1e-5 of their FP64 ones."""
    n = 128
    note("rotational vs divergence n=128")
    h = hit3d_amd.Hit3D(n, n_scalars=3)
    try:
        h.set_physics(1 / 800.0, flow_type=1, kfmax=2, famp=0.5, pe=(1.0, 1.0, 1.0),
                      scalar_type=(0, 0, 0))
        seeded(h)
        for c in range(3):
            h.copy(FIELDS, 4 + c, FIELDS, 1 + c)   # keep the initial field on the device
        out = {}
        for form, prec in ((0, 0), (1, 0), (1, 1), (2, 0), (2, 1)):
            for c in range(3):
                h.copy(FIELDS, 1 + c, FIELDS, 4 + c)
            h.set_nonlinear(form)
            fov = True
            for _ in range(5):
                _, _, fov, _ = h.step(1e-3, fov, True, precision=prec)
            out[(form, prec)] = np.stack([h.get(FIELDS, k) for k in (1, 2, 3)])
        assert rel(out[(1, 0)], out[(0, 0)]) < 1e-11
        assert rel(out[(2, 0)], out[(0, 0)]) < 1e-11
        for form in (1, 2):
            err = rel(out[(form, 1)], out[(form, 0)])
            note(f"mixed form {form} vs FP64 after 5 steps: {err:.2e}")
            assert err < 1e-5
    finally:
        h.close()


def test_rotational_form_limits():
    """The rotational form is an option of the fused 2/3-rule step: rejected
    with the phase shift, and a step that cannot run fused (active scalars)
    refuses it instead of silently using the divergence form."""
    h = hit3d_amd.Hit3D(16, dealias=1)
    try:
        with pytest.raises(hit3d_amd.Hit3dError, match="2/3"):
            h.set_nonlinear(1)
    finally:
        h.close()
    h = hit3d_amd.Hit3D(16, n_scalars=1)
    try:
        h.set_physics(0.01, pe=(0.01,), scalar_type=(0,))
        h.set_nonlinear(1)
        with pytest.raises(hit3d_amd.Hit3dError, match="rotational"):
            h.step(1e-3, True, True, int_scalars=True)
        with pytest.raises(hit3d_amd.Hit3dError, match="form"):
            h.set_nonlinear(3)
    finally:
        h.close()
