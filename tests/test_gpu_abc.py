"""Multi-step trajectories of the fused GPU steps against a closed form: the
ABC (Arnold-Beltrami-Childress) flow

    u = (a sin z + c cos y, b sin x + a cos z, c sin y + b cos x)

is a Beltrami field (curl u = u), so its nonlinear term -(u . grad) u =
-grad(|u|^2/2) is a gradient, nonzero but removed by the projection, and
every mode has |k| = 1.  The flow is an exact Navier-Stokes solution: after
m driver steps the fields are f0 y_m, with y_m from the Euler/AB2 recurrence
of dy/dt = -nu y (forced, KFMAX = 2: (FAMP/E - nu) y with the band energy
E = E0 y^2, all of it in shell 1).  The Courant number of step m is
y_{m-1} max(|u0| + |v0| + |w0|) dt/dx (rhs_velocity.f90:31-38).

The 2/3-rule step in its three nonlinear forms and the phase-shift step
(six and five products) at 64^3 over 30 steps, the benchmark's
configuration (512^3, forced, deviatoric form) and 1024^3 in FP64 and in
the mixed form.  Tolerances: fields 1e-12 of max |f0| in FP64, 1e-6 in the
mixed form (its single-RHS bound is 1e-5 of the RHS, and a step moves the
field by dt times the RHS); Courant number and band energy 1e-12 relative
(FP64); E(k) outside shell 1 below 1e-20 of E0 (FP64), 1e-12 (mixed).
"""
import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS, WRK

pytestmark = pytest.mark.gpu

A, B, C = 1.0, 0.8, 0.6
E0 = (A * A + B * B + C * C) / 2.0


def note(msg):
    print(msg, flush=True)


def abc_physical(n):
    """(u, v, w) of the ABC flow, each as a slot [z, y, x + 2 pad]."""
    x = 2 * np.pi * np.arange(n) / n
    s, c = np.sin(x), np.cos(x)
    out = []
    for expr in (A * s[:, None, None] + C * c[None, :, None],     # a sin z + c cos y
                 B * s[None, None, :] + A * c[:, None, None],     # b sin x + a cos z
                 C * s[None, :, None] + B * c[None, None, :]):    # c sin y + b cos x
        a = np.zeros((n, n, n + 2))
        a[:, :, :n] = expr
        out.append(a)
    return out


def load_abc(h, n):
    """Forward transforms of the physical ABC fields into the velocity
    slots; returns max(|u| + |v| + |w|) over the grid."""
    phys = abc_physical(n)
    m = np.abs(phys[0][:, :, :n]) + np.abs(phys[1][:, :, :n])
    m += np.abs(phys[2][:, :, :n])
    umax = float(m.max())
    del m
    for comp in range(3):
        h.put(WRK, 1, phys[comp])
        h.xFFT3d(1, 1)
        h.copy(FIELDS, comp + 1, WRK, 1)
    return umax


def abc_spectral(n):
    """The ABC fields' half spectra [ky, kz, kx re/im] written directly (the
    forward transform: sign -1, unnormalised): cos(k.x) has n^3/2 at k and at
    -k, sin(k.x) -i n^3/2 at k and +i n^3/2 at -k; a mode with kx = 0 stores
    both entries, one with kx = 1 only the first.  Used where the physical
    slots would not be worth their host memory (1024^3)."""
    h3 = float(n) ** 3 / 2.0
    out = [np.zeros((n, n, n + 2)) for _ in range(3)]

    def mode(a, k, re, im):
        kx, ky, kz = k
        a[ky % n, kz % n, 2 * kx] += re
        a[ky % n, kz % n, 2 * kx + 1] += im
        if kx == 0:
            a[-ky % n, -kz % n, 0] += re
            a[-ky % n, -kz % n, 1] -= im

    mode(out[0], (0, 0, 1), 0.0, -A * h3)   # a sin z
    mode(out[0], (0, 1, 0), C * h3, 0.0)    # c cos y
    mode(out[1], (1, 0, 0), 0.0, -B * h3)   # b sin x
    mode(out[1], (0, 0, 1), A * h3, 0.0)    # a cos z
    mode(out[2], (0, 1, 0), 0.0, -C * h3)   # c sin y
    mode(out[2], (1, 0, 0), B * h3, 0.0)    # b cos x
    return out


def amplitudes(steps, dt, nu, famp, forced):
    """y_0..y_steps: Euler first step, then AB2 (main.f90:124-176)."""
    rate = (lambda y: (famp / (E0 * y * y) - nu) * y) if forced else (lambda y: -nu * y)
    ys, r_prev = [1.0], None
    for _ in range(steps):
        r = rate(ys[-1])
        ys.append(ys[-1] + dt * (r if r_prev is None else 1.5 * r - 0.5 * r_prev))
        r_prev = r
    return ys


def run_abc(h, n, umax, steps, dt, nu, famp, forced, precision):
    """Steps the loaded ABC field (umax: its max |u| + |v| + |w|); checks the
    Courant number and band energy of every step; returns the amplitudes."""
    ys = amplitudes(steps, dt, nu, famp, forced)
    fov = True
    dx = 2 * np.pi / n
    tol = 1e-12 if precision == 0 else 1e-6
    for m in range(1, steps + 1):
        cr, en, fov, _ = h.step(dt, fov, True, precision=precision)
        expect = ys[m - 1] * umax * dt / dx
        assert abs(cr - expect) <= tol * expect, (m, cr, expect)
        if forced:
            assert abs(en - E0 * ys[m - 1] ** 2) <= tol * E0, (m, en)
    return ys


CASES_64 = [
    # dealias, nonlinear form, forced
    (0, 0, False),
    (0, 1, False),
    (0, 2, True),
    (1, 0, True),
    (1, 2, False),
]


@pytest.mark.parametrize("dealias,form,forced", CASES_64)
def test_abc_flow_64(dealias, form, forced):
    # dt keeps every kept mode inside AB2's stability region (nu k^2 dt <=
    # 0.48 for the phase shift's |k| <= 43.7), so roundoff does not grow
    n, steps, dt, nu, famp = 64, 30, 0.005, 0.05, 0.5
    note(f"abc flow n={n} dealias={dealias} form={form} forced={forced}")
    h = hit3d_amd.Hit3D(n, dealias=dealias)
    try:
        h.set_physics(nu, flow_type=1 if forced else 0, kfmax=2, famp=famp)
        h.set_nonlinear(form)
        umax = load_abc(h, n)
        f0 = np.stack([h.get(FIELDS, c) for c in (1, 2, 3)])
        # the directly written spectra (used at 1024^3) are the transforms
        assert np.abs(np.stack(abc_spectral(n)) - f0).max() < 1e-12 * np.abs(f0).max()
        ys = run_abc(h, n, umax, steps, dt, nu, famp, forced, 0)
        f = np.stack([h.get(FIELDS, c) for c in (1, 2, 3)])
        err = np.abs(f - ys[-1] * f0).max() / np.abs(f0).max()
        note(f"  y_{steps} = {ys[-1]:.6f}, field error {err:.2e}")
        assert err < 1e-12
        assert abs(ys[-1] - 1.0) > 1e-3
        e, _ = h.stats()
        assert abs(e[0] - E0 * ys[-1] ** 2) < 1e-12 * E0
        assert e[1:].sum() < 1e-20 * E0
        d = h.divergence()
        assert max(abs(d[0]), abs(d[1])) < 1e-10
    finally:
        h.set_nonlinear(0)
        h.close()


def test_abc_flow_512_benchmark_configuration():
    """512^3, forced (KFMAX = 2, FAMP = 0.5), RE = 800, deviatoric form: four
    steps in FP64 and in the mixed form, against the closed form."""
    n, steps, dt, nu, famp = 512, 4, 1e-3, 1 / 800.0, 0.5
    h = hit3d_amd.Hit3D(n, dealias=0)
    try:
        h.set_physics(nu, flow_type=1, kfmax=2, famp=famp)
        h.set_nonlinear(2)
        for precision in (0, 1):
            note(f"abc flow n=512 precision={precision}")
            umax = load_abc(h, n)
            w0 = h.get(FIELDS, 3)
            ys = run_abc(h, n, umax, steps, dt, nu, famp, True, precision)
            assert abs(ys[-1] - 1.0) > 1e-3
            e, _ = h.stats()
            tol = 1e-12 if precision == 0 else 1e-6
            assert abs(e[0] - E0 * ys[-1] ** 2) < tol * E0
            if precision == 0:
                assert e[1:].sum() < 1e-20 * E0
            d = h.divergence()
            assert max(abs(d[0]), abs(d[1])) < 1e-9
            # one component against its initial spectrum times y
            w = h.get(FIELDS, 3)
            err = np.abs(w - ys[-1] * w0).max() / np.abs(w0).max()
            note(f"  y_{steps} = {ys[-1]:.6f}, w field error {err:.2e}")
            assert err < tol
            del w, w0
    finally:
        h.set_nonlinear(0)
        h.close()


def test_abc_flow_1024():
    """1024^3 (its own kernels: half-length x pass, persistent column passes,
    paired half-width z/update blocks), forced, deviatoric form: four steps
    in FP64 and in the mixed form.  The spectra are written directly, and
    the check goes through the device statistics, so no slot crosses PCIe
    back: band energy per step E0 y^2, E(k) in shell 1 E0 y^2 and nothing
    elsewhere (an entry off by d relative would put d^2 E0 there)."""
    n, steps, dt, nu, famp = 1024, 4, 1e-3, 1 / 1600.0, 0.5
    h = hit3d_amd.Hit3D(n, dealias=0)
    try:
        h.set_physics(nu, flow_type=1, kfmax=2, famp=famp)
        h.set_nonlinear(2)
        for precision in (0, 1):
            note(f"abc flow n=1024 precision={precision}")
            for comp, a in enumerate(abc_spectral(n)):
                h.put(FIELDS, comp + 1, a)
                del a
            ys = amplitudes(steps, dt, nu, famp, True)
            tol = 1e-12 if precision == 0 else 1e-6
            fov = True
            for m in range(1, steps + 1):
                _, en, fov, _ = h.step(dt, fov, True, precision=precision)
                assert abs(en - E0 * ys[m - 1] ** 2) <= tol * E0, (m, en)
            e, _ = h.stats()
            note(f"  y_{steps} = {ys[-1]:.6f}, shell 1 {e[0] / (E0 * ys[-1] ** 2) - 1:.2e}, "
                 f"other shells {e[1:].sum() / E0:.2e}")
            assert abs(e[0] - E0 * ys[-1] ** 2) < tol * E0
            assert e[1:].sum() < (1e-20 if precision == 0 else 1e-12) * E0
            d = h.divergence()
            assert max(abs(d[0]), abs(d[1])) < 1e-9
    finally:
        h.set_nonlinear(0)
        h.close()
