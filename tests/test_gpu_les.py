"""GPU parity of the Smagorinsky LES path (les_model 1) against the oracle:
one RHS with the eddy-viscosity stress for velocities and scalars, and a short
forced trajectory.  Tolerance 1e-12 relative for one RHS."""
import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS, WRK

pytestmark = pytest.mark.gpu

RN1 = 67748937.0


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def rand_fields(oracle, n, seed, count):
    rng = np.random.default_rng(seed)
    f = np.zeros((count, n, n, n + 2))
    for c in range(count):
        a = np.zeros((n, n, n + 2))
        a[:, :, :n] = rng.uniform(-1, 1, (n, n, n))
        f[c] = oracle.fft_forward(a)
    f[:3] = oracle.pressure(f[:3], n)
    return oracle.dealias_all(f, oracle.kmax_of(n, 0), n)


def test_les_rhs_velocity_and_scalars(oracle):
    n = 32
    print("les rhs n=32", flush=True)
    nu = 1 / 800.0
    pe = (nu / 0.7,)
    h = hit3d_amd.Hit3D(n, n_scalars=1, dealias=0)
    h.set_physics(nu, pe=pe, scalar_type=(1,))
    h.set_les(1, 0.18)
    f = rand_fields(oracle, n, 70, 4)
    for c in range(4):
        h.put(FIELDS, c + 1, f[c])
    h.rhs_scalars(True)
    tv = oracle.turb_visc_smag(f, n, 0.18)
    rs = oracle.rhs_scalars(f, 1, pe, (1,), 0, n)
    rs = oracle.les_rhs_scalar(f[3], rs[0], tv, n)
    assert rel(h.get(WRK, 4), rs) < 1e-12
    h.rhs_velocity(0.001)
    rv, _ = oracle.rhs_velocity(f, nu, 0, n)
    rv = oracle.les_rhs_velocity(f, rv, tv, n)
    for c in range(3):
        assert rel(h.get(WRK, c + 1), rv[c]) < 1e-12
    # nu_t as io_write_4 writes it (io_write_4.f90:73-79): float32 of turb_visc
    nut = h.turb_visc()
    assert nut.dtype == np.float32 and nut.shape == (n, n, n)
    assert np.array_equal(nut, tv.astype(np.float32)) or rel(nut, tv) < 1e-7
    h.close()


def test_les_trajectory(oracle):
    n = 32
    print("les trajectory n=32", flush=True)
    cfg = oracle.Config(n=n, nu=1 / 800.0, dt=0.002, dealias=0, flow_type=1, kfmax=2, famp=0.5,
                        les_model=1, c_smag=0.18)
    h = hit3d_amd.Hit3D(n, dealias=0)
    h.set_physics(cfg.nu, flow_type=1, kfmax=2, famp=0.5)
    h.set_les(1, 0.18)
    h.init_velocity(RN1, 2, 4, 3.0)
    h.dealias_all()
    f0 = np.stack([h.get(FIELDS, c) for c in (1, 2, 3)])
    st = oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt)
    fov = True
    for it in range(5):
        c, e, fov, _ = h.step(cfg.dt, fov, True)
        st = oracle.step(cfg, st)
        assert abs(c - st.courant) < 1e-12 * st.courant
        assert abs(e - st.energy) < 1e-12 * st.energy
    assert rel(np.stack([h.get(FIELDS, c) for c in (1, 2, 3)]), st.fields) < 1e-11
    h.close()


def test_les_rejects_phase_shift():
    h = hit3d_amd.Hit3D(16, dealias=1)
    with pytest.raises(hit3d_amd.Hit3dError, match="2/3"):
        h.set_les(1)
    # no LES: no turbulent viscosity to write
    with pytest.raises(hit3d_amd.Hit3dError, match="turb_visc"):
        h.turb_visc()
    h.close()
