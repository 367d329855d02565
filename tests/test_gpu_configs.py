"""The measured configurations' physics under GPU test at sizes the oracle
finishes in seconds, through the C ABI.

* 32^3 decaying turbulence with the 2/3 rule, RE = 100, fixed DT = 0.01,
  for the configuration's full 200 steps (fused step, graph replay), with
  E(k), energy and dissipation compared with the oracle's after every step.
* A miniature of the 1024^3 configuration at 32^3: forced, RE = 1600,
  phase shift, three type-0 scalars (Sc 0.5, 1, 2) with the mean-gradient
  source, variable DT, and CINT particles moved every step with the
  velocity the step starts from (the parts task of main.f90:235-247).
  Fields, Courant number, band energy and dt against the oracle's
  trajectory; every particle move bit for bit against the oracle's move
  of the same particles through the same velocity.
"""
import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS, WRK

pytestmark = pytest.mark.gpu

RN1 = 67748937.0


def note(msg):
    print(msg, flush=True)


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def test_decaying_32_full_200_steps(oracle):
    """32^3, flow_type 0, RE 100, DT = 0.01 fixed, 2/3 rule, 200 steps from
    the seeded k^4 exp(-2 (k/3)^2) field.  Per step: E(k) within 1e-10,
    energy and dissipation within 1e-11 of the oracle's (the two
    trajectories start bit-identical and drift apart only by the transforms'
    rounding); the final fields within 1e-10."""
    n, steps = 32, 200
    cfg = oracle.Config(n=n, nu=0.01, dt=0.01, dealias=0, flow_type=0)
    note("decaying 32^3, 200 steps")
    h = hit3d_amd.Hit3D(n, dealias=0)
    try:
        h.set_physics(cfg.nu, flow_type=0)
        h.init_velocity(RN1, 2, 4, 3.0)
        h.dealias_all()
        f0 = np.stack([h.get(FIELDS, c) for c in (1, 2, 3)])
        st = oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt)
        fov = True
        worst_e = worst_s = 0.0
        for it in range(steps):
            c, _, fov, _ = h.step(cfg.dt, fov, True)
            st = oracle.step(cfg, st)
            assert abs(c - st.courant) <= 1e-11 * st.courant, it
            e, _ = h.stats()
            ref = oracle.stats(st.fields, n, cfg.kmax, cfg.nu, 1 / cfg.nu)
            worst_e = max(worst_e, rel(e, ref["e_spec"]))
            energy = float(np.sum(e))
            eps = float(np.sum(e * np.arange(1, cfg.kmax + 1) ** 2 * 2.0 * cfg.nu))
            worst_s = max(worst_s, abs(energy - ref["energy"]) / ref["energy"],
                          abs(eps - ref["eps_v"]) / ref["eps_v"])
            assert worst_e < 1e-10 and worst_s < 1e-11, (it, worst_e, worst_s)
        note(f"200 steps: max E(k) rel err {worst_e:.2e}, energy/eps {worst_s:.2e}")
        got = np.stack([h.get(FIELDS, c) for c in (1, 2, 3)])
        assert rel(got, st.fields) < 1e-10
        # the run decays: energy after 200 steps is below the initial energy
        e0 = oracle.stats(f0, n, cfg.kmax, cfg.nu, 1 / cfg.nu)["energy"]
        assert ref["energy"] < 0.9 * e0
    finally:
        h.close()


def test_phase_shift_scalars_particles_32(oracle):
    """32^3 miniature of the 1024^3 configuration (see module docstring):
    10 steps.  Fields within 1e-11 of the oracle's trajectory, scalars
    included; Courant number, band energy within 1e-12 and dt equal per
    step; each move of the 4096 CINT particles bit-exact against the
    oracle's move from the same positions through the same velocity, and
    the particle trajectory within 1e-10 grid units of the oracle's."""
    n, steps, npart = 32, 10, 4096
    sc = (0.5, 1.0, 2.0)
    cfg = oracle.Config(n=n, nu=1 / 1600.0, dt=0.001, dealias=1, flow_type=1, kfmax=2,
                        famp=0.5, variable_dt=True, n_scalars=3, sc=sc, scalar_type=(0, 0, 0))
    note("phase shift + 3 scalars + CINT particles 32^3")
    h = hit3d_amd.Hit3D(n, n_scalars=3, dealias=1)
    try:
        h.set_physics(cfg.nu, flow_type=1, kfmax=2, famp=0.5, pe=cfg.pe, scalar_type=(0, 0, 0))
        h.init_velocity(RN1, 2, 4, 3.0)
        h.dealias_all()
        f0 = np.zeros((6, n, n, n + 2))
        for c in range(3):
            f0[c] = h.get(FIELDS, c + 1)
        st = oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt)
        rng = np.random.default_rng(2024)
        pos = rng.uniform(1.0, n + 1.0, (npart, 3))
        uvw = np.full((npart, 3), -999.0)
        h.particles_set(pos)
        fov = fos = True
        dt = cfg.dt
        for it in range(steps):
            # parts task: the velocity this step starts from, then the move
            h.fields_to_parts()
            vel = np.stack([h.get(WRK, c) for c in (1, 2, 3)])
            ovel = np.stack([oracle.fft_inverse(st.fields[c]) for c in range(3)])
            assert rel(vel[:, :, :, :n], ovel[:, :, :, :n]) < 1e-11, it
            gx0, gu0 = h.particles_get()
            h.particles_move(dt, 1)
            gx, gu = h.particles_get()
            ox, ou = oracle.particles_move(vel, gx0, gu0, dt, n, 1)
            assert np.array_equal(gx, ox) and np.array_equal(gu, ou), it
            pos, uvw = oracle.particles_move(ovel, pos, uvw, dt, n, 1)
            d = np.abs(gx - pos)
            d = np.minimum(d, n - d)   # periodic seam
            assert d.max() < 1e-10, it
            # hydro step
            c, e, fov, fos = h.step(dt, fov, fos, int_scalars=True)
            st = oracle.step(cfg, st)
            assert abs(c - st.courant) <= 1e-12 * st.courant, it
            assert abs(e - st.energy) <= 1e-12 * st.energy, it
            dt = oracle.my_dt(dt, c, True, sc)
            assert dt == st.dt, it
        for c in range(6):
            assert rel(h.get(FIELDS, c + 1), st.fields[c]) < 1e-11, c
        # the scalars were driven by the mean gradient and carry variance
        assert np.abs(st.fields[3:]).max() > 0.0
        ia = oracle.ialias(n)
        for c in range(6):
            assert np.all(h.get(FIELDS, c + 1)[ia >= 2] == 0.0), c
    finally:
        h.close()
