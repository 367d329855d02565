"""The fused velocity RHS of the phase-shift step (step_ps.hip: products
and half-spectrum transforms in one x pass, both halves' assembly, the
shift back, the mask, forcing, AB2 update and projection in the z passes)
against the unfused sequence of the reference's subroutines on the same
library, and against the oracle.

* Fused vs unfused, 3 driver steps (Euler, AB2) from the seeded field,
  with and without scalars (the scalars' fused RHS and update too), n =
  32..256: fields and rhs_old, scalars included, within 1e-12; Courant
  number and band energy within 1e-13.
* Fused vs the oracle at 32^3 and 64^3: a forced phase-shift trajectory
  within 1e-11, every entry with ialias >= 2 exactly zero.
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


def _run(h, n, nf, steps, fused, scal, f0):
    for c in range(nf):
        h.put(FIELDS, c + 1, f0[c])
    h.set_fusion(fused)
    fov = fos = True
    out = []
    for _ in range(steps):
        c, e, fov, fos = h.step(2e-3, fov, fos, int_scalars=scal)
        out.append((c, e))
    f = np.stack([h.get(FIELDS, c + 1) for c in range(nf)])
    r = np.stack([h.get(RHS_OLD, c + 1) for c in range(nf)])
    return out, f, r


@pytest.mark.parametrize("n,ns", [(32, 0), (64, 2), (128, 0), (256, 2)])
def test_fused_phase_shift_matches_unfused(n, ns):
    note(f"fused phase shift n={n} scalars={ns}")
    h = hit3d_amd.Hit3D(n, n_scalars=ns, dealias=1)
    try:
        nu = 1 / 400.0
        h.set_physics(nu, flow_type=1, kfmax=2, famp=0.5, pe=tuple(nu / s for s in (0.7, 1.0)[:ns]),
                      scalar_type=(0,) * ns)
        h.init_velocity(RN1, 2, 4, 3.0)
        h.dealias_all()
        nf = 3 + ns
        f0 = np.stack([h.get(FIELDS, c + 1) for c in range(nf)])
        try:
            a = _run(h, n, nf, 3, True, ns > 0, f0)
            b = _run(h, n, nf, 3, False, ns > 0, f0)
        finally:
            h.set_fusion(True)
        for (c1, e1), (c2, e2) in zip(a[0], b[0]):
            assert abs(c1 - c2) <= 1e-13 * c2 and abs(e1 - e2) <= 1e-13 * e2
        note(f"fields {rel(a[1], b[1]):.2e} rhs_old {rel(a[2], b[2]):.2e}")
        assert rel(a[1], b[1]) < 1e-12 and rel(a[2], b[2]) < 1e-12
    finally:
        h.close()


@pytest.mark.parametrize("n", [32, 64])
def test_fused_phase_shift_matches_oracle(oracle, n):
    note(f"fused phase shift vs oracle n={n}")
    cfg = oracle.Config(n=n, nu=1 / 400.0, dt=0.002, dealias=1, flow_type=1, kfmax=2, famp=0.5,
                        variable_dt=True)
    h = hit3d_amd.Hit3D(n, dealias=1)
    try:
        h.set_physics(cfg.nu, flow_type=1, kfmax=2, famp=0.5)
        h.init_velocity(RN1, 2, 4, 3.0)
        h.dealias_all()
        f0 = np.stack([h.get(FIELDS, c) for c in (1, 2, 3)])
        st = oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt)
        fov, dt = True, cfg.dt
        for it in range(6):
            c, e, fov, _ = h.step(dt, fov, True)
            st = oracle.step(cfg, st)
            assert abs(c - st.courant) < 1e-12 * st.courant, it
            assert abs(e - st.energy) < 1e-12 * st.energy, it
            dt = oracle.my_dt(dt, c)
            assert dt == st.dt
        got = np.stack([h.get(FIELDS, c) for c in (1, 2, 3)])
        assert rel(got, st.fields) < 1e-11
        ia = oracle.ialias(n)
        for c in range(3):
            assert np.all(got[c][ia >= 2] == 0.0)
    finally:
        h.close()


@pytest.mark.parametrize("n,ns", [(32, 2), (64, 0)])
def test_phase_shift_five_product_form(n, ns):
    """hit3d_set_nonlinear(2) in the fused phase-shift step: the five
    products uu - ww, uv, uw, vv - ww, vw instead of the six u_i u_j.  Their
    divergence differs by the gradient of ww, masked where ialias >= 2 like
    the rest of the RHS, which the projection removes: after 3 driver
    steps the fields (scalars included) equal the six-product form's to
    rounding, the Courant numbers and band energies too, and the difference
    of the two rhs_old is parallel to k at every wavevector (k x dr = 0)
    and not zero."""
    note(f"phase shift five-product form n={n} scalars={ns}")
    h = hit3d_amd.Hit3D(n, n_scalars=ns, dealias=1)
    try:
        nu = 1 / 400.0
        h.set_physics(nu, flow_type=1, kfmax=2, famp=0.5, pe=tuple(nu / s for s in (0.7, 1.0)[:ns]),
                      scalar_type=(0,) * ns)
        h.init_velocity(RN1, 2, 4, 3.0)
        h.dealias_all()
        nf = 3 + ns
        f0 = np.stack([h.get(FIELDS, c + 1) for c in range(nf)])
        a = _run(h, n, nf, 3, True, ns > 0, f0)
        h.set_nonlinear(2)
        try:
            b = _run(h, n, nf, 3, True, ns > 0, f0)
        finally:
            h.set_nonlinear(0)
        for (c1, e1), (c2, e2) in zip(a[0], b[0]):
            assert abs(c1 - c2) <= 1e-13 * c2 and abs(e1 - e2) <= 1e-13 * e2
        note(f"fields {rel(b[1], a[1]):.2e}")
        assert rel(b[1], a[1]) < 1e-12
        # rhs_old difference: a gradient i k phi in the (ky, kz, kx re/im) layout
        d = b[2][:3] - a[2][:3]
        kx = np.arange(n // 2 + 1)[None, None, :]
        kk = np.array([j if 2 * j <= n else j - n for j in range(n)], float)  # x_fftw.f90:280-296
        ky, kz = kk[:, None, None], kk[None, :, None]
        dc = d[:, :, :, 0::2] + 1j * d[:, :, :, 1::2]
        cross = [ky * dc[2] - kz * dc[1], kz * dc[0] - kx * dc[2], kx * dc[1] - ky * dc[0]]
        kmax = np.sqrt(3.0) * n / 2
        scale = np.abs(dc).max()
        assert scale > 1e-8 * np.abs(a[2][:3]).max()
        assert max(np.abs(c).max() for c in cross) <= 1e-11 * kmax * scale
    finally:
        h.close()
