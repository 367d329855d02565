"""Pins of the oracle's phase-shift de-aliasing, the seeded initial field's
orientation and the driver-step composition, each against something fixed
outside the oracle's own code: exact products on a zero-padded grid, known
values of the Box-Muller formula, the analytic curl of single modes, and the
closed-form evolution of a forced shear wave.

Each test names the mistake it would catch.
"""
import numpy as np
import pytest


def _modes_field(n, modes, m=None):
    """Physical real field sum_k a cos(k.x) + b sin(k.x) sampled on an m^3
    grid (m defaults to n), returned as [z, y, x] of shape (m, m, m)."""
    m = m or n
    x = np.arange(m) * (2 * np.pi / m)
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    out = np.zeros((m, m, m))
    for (kx, ky, kz), a, b in modes:
        ph = kx * xx + ky * yy + kz * zz
        out += a * np.cos(ph) + b * np.sin(ph)
    return out


def _pad(p, n):
    out = np.zeros((n, n, n + 2))
    out[:, :, :n] = p
    return out


def _random_modes(rng, kxs, kys, kzs, count):
    out = []
    for _ in range(count):
        k = (int(rng.choice(kxs)), int(rng.choice(kys)), int(rng.choice(kzs)))
        out.append((k, float(rng.standard_normal()), float(rng.standard_normal())))
    return out


def _exact_coefficients(n, prod_2n):
    """Fourier coefficients of a product sampled on the (2n)^3 grid (exact
    for band-limited factors), scaled to the unnormalised n^3 convention and
    laid out as the n-grid half spectrum [ky, kz, kx] (complex).  Modes with
    |k_i| >= n/2 are dropped: the test fields keep products off them."""
    m = 2 * n
    c = np.fft.rfftn(prod_2n, axes=(0, 1, 2)) * (float(n) ** 3 / float(m) ** 3)  # [kz, ky, kx]
    out = np.zeros((n, n, n // 2 + 1), complex)
    ks = [k if k <= n // 2 else k - n for k in range(n)]
    for iy, ky in enumerate(ks):
        for iz, kz in enumerate(ks):
            if abs(ky) >= n // 2 or abs(kz) >= n // 2:
                continue
            out[iy, iz, :n // 2] = c[kz % m, ky % m, :n // 2]
    return out


# single aliasing in x only: |kx| in {6, 7} at n = 16 (n/3 = 5.33), sums
# 12..14 wrap to -4..-2, never reach the Nyquist column; |ky|, |kz| <= 1
N = 16
KXS = [6, 7, -6, -7, 0]
KYS = [-1, 0, 1]
KZS = [-1, 0, 1]


def _velocity_modes(seed):
    rng = np.random.default_rng(seed)
    return [_random_modes(rng, KXS, KYS, KZS, 5) for _ in range(3)]


@pytest.mark.parametrize("seed", [1, 2])
def test_phase_shift_velocity_rhs_equals_exact_products(oracle, seed):
    """rhs_velocity.f90:150-410 with fields whose products alias once (in
    x): the half-cell shift of :248-249 cancels every such alias, so the
    phase-shift RHS equals -i k_j (u_i u_j)^ - nu k^2 u_i with the products
    taken exactly (zero-padded 2n grid) on every entry with ialias < 2, and
    is zero where ialias >= 2.  The 2/3-rule RHS keeps the aliases.
    Catches a shift of dx instead of dx/2 (no cancellation), a wrong mask
    threshold (ialias > 0 / >= 1 zeroes kept modes), a missing back-shift
    or a wrong sign of the shift."""
    n, nu = N, 0.05
    modes = _velocity_modes(seed)
    U = np.stack([oracle.fft_forward(_pad(_modes_field(n, md), n)) for md in modes])
    rhs, _ = oracle.rhs_velocity(U, nu, 1, n)
    u2 = [_modes_field(n, md, 2 * n) for md in modes]
    P = {}
    for i in range(3):
        for j in range(i, 3):
            P[(i, j)] = P[(j, i)] = _exact_coefficients(n, u2[i] * u2[j])
    kx, ky, kz = oracle._kgrid(n)
    k = (kx, ky, kz)
    k2 = (kx ** 2 + ky ** 2) + kz ** 2
    ia = oracle.ialias(n)[:, :, 0::2]
    assert (ia == 1).sum() > 0
    worst = 0.0
    for i in range(3):
        Ui = oracle.to_complex(U[i])
        ex = -1j * sum(k[j] * P[(i, j)] for j in range(3)) - nu * k2 * Ui
        ex = np.where(ia >= 2, 0.0, ex)
        got = oracle.to_complex(rhs[i])
        scale = np.abs(ex).max()
        worst = max(worst, np.abs(got - ex).max() / scale)
        # entries with one component above n/3 carry content
        assert np.abs(np.where(ia == 1, ex, 0.0)).max() > 0.1 * scale
        assert np.all(got[ia >= 2] == 0.0)
    assert worst < 1e-13
    # the 2/3 rule (rhs_velocity.f90:53-144) keeps the aliases on ialias == 0
    rhs23, _ = oracle.rhs_velocity(U, nu, 0, n)
    diff = 0.0
    for i in range(3):
        Ui = oracle.to_complex(U[i])
        ex = -1j * sum(k[j] * P[(i, j)] for j in range(3)) - nu * k2 * Ui
        got = oracle.to_complex(rhs23[i])
        diff = max(diff, np.abs(np.where(ia == 0, got - ex, 0.0)).max() / np.abs(ex).max())
    assert diff > 1e-2


def test_phase_shift_velocity_masks_double_alias_entries(oracle):
    """Fields carrying modes with two components above n/3: every entry
    with ialias >= 2 of the phase-shift RHS is exactly zero (the viscous
    term too), and entries with ialias == 1 are not (rhs_velocity.f90:215,
    :387).  Catches a final mask of ialias >= 3 or ialias >= 1."""
    n = N
    rng = np.random.default_rng(7)
    modes = [_random_modes(rng, [6, 7, 0], [6, -6, 1], [0, 1, 7], 6) for _ in range(3)]
    U = np.stack([oracle.fft_forward(_pad(_modes_field(n, md), n)) for md in modes])
    rhs, _ = oracle.rhs_velocity(U, 0.1, 1, n)
    ia = oracle.ialias(n)
    for i in range(3):
        assert np.all(rhs[i][ia >= 2] == 0.0)
        assert np.abs(rhs[i][ia == 1]).max() > 0.0


@pytest.mark.parametrize("seed", [3, 4])
def test_phase_shift_scalar_rhs_equals_exact_products(oracle, seed):
    """rhs_scalars.f90:144-271 and the gradient source :387-389 with a
    velocity and a scalar whose products alias once (in x): the RHS equals
    -i k_j (phi u_j)^ - pe k^2 phi - u_1 (type 0) with exact products on
    every entry with ialias < 2, and -u_1 where ialias >= 2.  The 2/3-rule
    RHS keeps the aliases.  Catches the scalar path's shift (:152-153),
    its two masks (:200, :257) and a sign of the back-shift."""
    n, pe = N, 0.07
    modes = _velocity_modes(seed)
    rng = np.random.default_rng(seed + 100)
    smodes = _random_modes(rng, KXS, KYS, KZS, 5)
    U = [oracle.fft_forward(_pad(_modes_field(n, md), n)) for md in modes]
    S = oracle.fft_forward(_pad(_modes_field(n, smodes), n))
    f = np.stack(U + [S])
    r = oracle.rhs_scalars(f, 1, (pe,), (0,), 1, n)[0]
    phi2 = _modes_field(n, smodes, 2 * n)
    Q = [_exact_coefficients(n, phi2 * _modes_field(n, md, 2 * n)) for md in modes]
    kx, ky, kz = oracle._kgrid(n)
    k2 = (kx ** 2 + ky ** 2) + kz ** 2
    ia = oracle.ialias(n)[:, :, 0::2]
    Sc = oracle.to_complex(S)
    U1 = oracle.to_complex(U[0])
    ex = -1j * ((kx * Q[0] + ky * Q[1]) + kz * Q[2]) - pe * k2 * Sc
    ex = np.where(ia >= 2, 0.0, ex) - U1
    got = oracle.to_complex(r)
    assert np.abs(got - ex).max() < 1e-13 * np.abs(ex).max()
    assert np.abs(np.where(ia == 1, ex + U1, 0.0)).max() > 0.1 * np.abs(ex).max()
    r23 = oracle.rhs_scalars(f, 1, (pe,), (0,), 0, n)[0]
    diff = np.abs(np.where(ia == 0, oracle.to_complex(r23) - ex, 0.0)).max()
    assert diff > 1e-2 * np.abs(ex).max()


# ------------------------------------------------------------ Box-Muller, curl

def test_box_muller_known_values(oracle):
    """init_velocity.f90:113, g = sqrt(-2 log a) sin(2 pi b): a = e^(-1/2)
    gives the factor 1, so b = 1/4 -> 1, b = 1/2 -> 0, b = 3/4 -> -1,
    b = 1/12 -> 1/2.  Catches sin/cos swapped, a missing factor 2 in either
    place, or the two slots swapped (a and b exchanged)."""
    a = np.full(4, np.exp(-0.5))
    b = np.array([0.25, 0.5, 0.75, 1.0 / 12.0])
    g = oracle.box_muller(a, b)
    assert abs(g[0] - 1.0) < 1e-15
    assert abs(g[1]) < 1e-15
    assert abs(g[2] + 1.0) < 1e-15
    assert abs(g[3] - 0.5) < 1e-15
    # a = e^(-2) gives the factor 2
    g2 = oracle.box_muller(np.array([np.exp(-2.0)]), np.array([0.25]))
    assert abs(g2[0] - 2.0) < 1e-15


def test_init_velocity_curl_single_modes(oracle):
    """The post-RNG stage of init_velocity.f90:116-237 on Gaussian slots
    g1 = sum_m cos mz, g2 = sum_m cos mx, g3 = sum_m cos my (m = 1..kmax):
    the curl i k x W (:123-147) is (dg3/dy - dg2/dz, dg1/dz - dg3/dx,
    dg2/dx - dg1/dy) = -sum_m m (sin my, sin mz, sin mx), energy 3 m^2/4 in
    shell m; the rescale (:153-237) brings shell m to the normalised
    target, amplitude sqrt(4 target_m / 3).  Catches a curl of the wrong
    sign, components permuted, or a rescale by E/target instead of
    target/E."""
    n, kmax = 16, 5
    ms = range(1, kmax + 1)
    g = np.stack([_pad(_modes_field(n, [((0, 0, m), 1.0, 0.0) for m in ms]), n),
                  _pad(_modes_field(n, [((m, 0, 0), 1.0, 0.0) for m in ms]), n),
                  _pad(_modes_field(n, [((0, m, 0), 1.0, 0.0) for m in ms]), n)])
    f, e0 = oracle.init_velocity_from_gaussian(g, n, kmax, 2, 3.0, 4.0)
    tgt = oracle.target_spectrum(kmax, 2, 3.0, 4.0)
    assert np.allclose(e0, [3.0 * m * m / 4.0 for m in ms], rtol=1e-13, atol=0)
    amp = [np.sqrt(4.0 * tgt[m - 1] / 3.0) for m in ms]
    ex = [_modes_field(n, [((0, m, 0), 0.0, -a) for m, a in zip(ms, amp)]),
          _modes_field(n, [((0, 0, m), 0.0, -a) for m, a in zip(ms, amp)]),
          _modes_field(n, [((m, 0, 0), 0.0, -a) for m, a in zip(ms, amp)])]
    for c in range(3):
        got = oracle.fft_inverse(f[c])[:, :, :n]
        assert np.abs(got - ex[c]).max() < 1e-13


# ------------------------------------------------------------ step composition

def _shear_wave(oracle, n, a):
    """w = a sin x (u = v = 0): divergence-free, its nonlinear term is zero
    in divergence form (d_x(u w) = 0, d_z(w w) = 0) and its only entry is
    kx = 1, in forcing shell 1."""
    z = np.zeros((n, n, n + 2))
    return np.stack([z, z.copy(),
                     oracle.fft_forward(_pad(_modes_field(n, [((1, 0, 0), 0.0, a)]), n))])


def test_step_forced_shear_wave_closed_form(oracle):
    """main.f90:124-176 on a forced shear wave w = A sin x, KFMAX = 2: the
    RHS is (FAMP/E - nu) w-hat with E = A^2/4 the band energy, so Euler
    gives A1 = A0 (1 + dt (FAMP/E0 - nu)) and AB2 then gives
    A2 = A1 + dt (1.5 r1 - 0.5 r0).  Catches the forcing sign or factor
    (m_force.f90:163-173), AB2 weights or order, the Euler first step,
    and forcing applied after the update instead of to the RHS."""
    n, nu, famp, dt, a0 = 16, 0.02, 0.5, 0.01, 0.8
    cfg = oracle.Config(n=n, nu=nu, dt=dt, dealias=0, flow_type=1, kfmax=2, famp=famp)
    f0 = _shear_wave(oracle, n, a0)
    st = oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=dt)
    st1 = oracle.step(cfg, st)
    e0 = a0 * a0 / 4.0
    r0 = (famp / e0 - nu) * a0
    a1 = a0 + dt * r0
    assert abs(st1.energy - e0) < 1e-15
    e1 = a1 * a1 / 4.0
    r1 = (famp / e1 - nu) * a1
    a2 = a1 + dt * (1.5 * r1 - 0.5 * r0)
    st2 = oracle.step(cfg, st1)
    assert abs(st2.energy - e1) < 1e-14
    for s, a in ((st1, a1), (st2, a2)):
        w = oracle.fft_inverse(s.fields[2])[:, :, :n]
        assert np.abs(w - _modes_field(n, [((1, 0, 0), 0.0, a)])).max() < 1e-13
        assert np.abs(s.fields[:2]).max() < 1e-13
    # Courant number of w = A sin x: max |w| dt / dx (rhs_velocity.f90:31-38)
    assert abs(st1.courant - a0 * dt / (2 * np.pi / n)) < 1e-14
    # decaying (flow_type 0): pure viscous decay, A1 = A0 (1 - nu dt)
    cfg0 = oracle.Config(n=n, nu=nu, dt=dt, dealias=0, flow_type=0)
    s0 = oracle.step(cfg0, oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=dt))
    w = oracle.fft_inverse(s0.fields[2])[:, :, :n]
    assert np.abs(w - _modes_field(n, [((1, 0, 0), 0.0, a0 * (1 - nu * dt))])).max() < 1e-14


def test_step_projects_after_the_update(oracle):
    """main.f90:150-167 projects the updated velocity, not the RHS alone or
    the old field: from a field whose RHS has a gradient part, the stepped
    field is solenoidal to roundoff, and equals the projection of the
    unprojected update.  A projection before the update leaves dt times
    the RHS's gradient part in the field."""
    n, dt = 16, 0.01
    f = np.stack([oracle.fft_forward(_pad(_modes_field(n, [((1, 2, 0), 0.0, 1.0)]), n)),
                  oracle.fft_forward(_pad(_modes_field(n, [((0, 0, 3), 1.0, 0.0)]), n)),
                  oracle.fft_forward(_pad(_modes_field(n, [((2, 0, 0), 0.0, 1.0)]), n))])
    cfg = oracle.Config(n=n, nu=0.0, dt=dt, dealias=0, flow_type=0)
    st = oracle.step(cfg, oracle.State(fields=f, rhs_old=np.zeros_like(f), dt=dt))
    kx, ky, kz = oracle._kgrid(n)
    C = [oracle.to_complex(st.fields[c]) for c in range(3)]
    assert np.abs((kx * C[0] + ky * C[1]) + kz * C[2]).max() < 1e-10
    rhs, _ = oracle.rhs_velocity(f, 0.0, 0, n)
    D = [oracle.to_complex(rhs[c]) for c in range(3)]
    grad = np.abs((kx * D[0] + ky * D[1]) + kz * D[2]).max()
    assert grad > 1.0   # the RHS is not solenoidal
    # the step result differs from "project the old field, then add dt*rhs"
    wrong = oracle.pressure(f, n) + dt * rhs
    assert np.abs(st.fields - wrong).max() > 1e-3 * dt * grad


# ------------------------------------------------------------ rotational form

def test_rotational_rhs_shear_closed_form(oracle):
    """u = (sin y, 0, 0): omega = (0, 0, -cos y), u x omega = (0, sin y cos y,
    0) = grad(|u|^2/2) (u . grad u = 0), so the rotational RHS is that
    gradient minus nu k^2 u, while the divergence form has no nonlinear
    term.  Catches a sign of omega, of the cross product, or a permuted
    component."""
    n, nu = 16, 0.1
    z = np.zeros((n, n, n + 2))
    f = np.stack([oracle.fft_forward(_pad(_modes_field(n, [((0, 1, 0), 0.0, 1.0)]), n)), z, z])
    rhs, _ = oracle.rhs_velocity_rotational(f, nu, n)
    got = [oracle.fft_inverse(r)[:, :, :n] for r in rhs]
    ex = [_modes_field(n, [((0, 1, 0), 0.0, -nu)]),
          _modes_field(n, [((0, 2, 0), 0.0, 0.5)]),     # sin y cos y = sin(2y)/2
          np.zeros((n, n, n))]
    for g, e in zip(got, ex):
        assert np.abs(g - e).max() < 1e-13


def test_rotational_rhs_projects_to_divergence_form(oracle):
    """For band-limited fields the 2/3 rule is exact for both forms, and
    u x omega = -d_j(u_i u_j) + grad(|u|^2/2) for solenoidal u: the
    projected RHS of the two forms agree to rounding, and the unprojected
    ones differ by a gradient (k x (r_rot - r_div) = 0)."""
    n, nu = 32, 0.02
    rng = np.random.default_rng(5)
    modes = [_random_modes(rng, [0, 1, 2, 3, -2], [0, 1, -1, 3], [0, 2, -1, -3], 6)
             for _ in range(3)]
    U = np.stack([oracle.fft_forward(_pad(_modes_field(n, md), n)) for md in modes])
    U = oracle.pressure(U, n)
    r_div, _ = oracle.rhs_velocity(U, nu, 0, n)
    r_rot, _ = oracle.rhs_velocity_rotational(U, nu, n)
    pd, pr = oracle.pressure(r_div, n), oracle.pressure(r_rot, n)
    assert np.abs(pr - pd).max() < 1e-12 * np.abs(pd).max()
    d = [oracle.to_complex(r_rot[c] - r_div[c]) for c in range(3)]
    assert np.abs(d[0]).max() > 1e-3 * np.abs(pd).max()   # a gradient part exists
    kx, ky, kz = oracle._kgrid(n)
    curl = [ky * d[2] - kz * d[1], kz * d[0] - kx * d[2], kx * d[1] - ky * d[0]]
    assert max(np.abs(c).max() for c in curl) < 1e-11 * np.abs(pd).max()


def test_rotational_step_equals_divergence_step(oracle):
    """Driver steps (Euler then AB2, forcing, projection) with the two forms
    from one solenoidal band-limited field give the same fields to rounding:
    the gradient part of the rotational RHS, kept in rhs_old, is removed by
    the projection after every update."""
    n = 32
    rng = np.random.default_rng(8)
    modes = [_random_modes(rng, [0, 1, 2, -1], [0, 1, -2, 2], [0, 1, 2, -1], 6)
             for _ in range(3)]
    U = np.stack([oracle.fft_forward(_pad(_modes_field(n, md), n)) for md in modes])
    U = oracle.dealias_all(oracle.pressure(U, n), n // 3, n)
    out = {}
    for rot in (False, True):
        cfg = oracle.Config(n=n, nu=0.01, dt=0.002, dealias=0, flow_type=1, kfmax=2, famp=0.5,
                            rotational=rot)
        st = oracle.State(fields=U.copy(), rhs_old=np.zeros_like(U), dt=cfg.dt)
        for _ in range(3):
            st = oracle.step(cfg, st)
        out[rot] = st
    a, b = out[False], out[True]
    assert np.abs(b.fields - a.fields).max() < 1e-12 * np.abs(a.fields).max()
    assert abs(b.courant - a.courant) < 1e-12 * a.courant
    assert abs(b.energy - a.energy) < 1e-12 * a.energy


# ------------------------------------------------------------ deviatoric form

def test_deviatoric_rhs_shear_closed_form(oracle):
    """u = (sin y, 0, 0): the divergence form's nonlinear term is zero, the
    trace-free tensor's is grad(|u|^2/3) = (0, sin(2y)/3, 0).  Catches a
    trace removed with the wrong factor or on the wrong components."""
    n, nu = 16, 0.1
    z = np.zeros((n, n, n + 2))
    f = np.stack([oracle.fft_forward(_pad(_modes_field(n, [((0, 1, 0), 0.0, 1.0)]), n)), z, z])
    rhs, _ = oracle.rhs_velocity_deviatoric(f, nu, n)
    got = [oracle.fft_inverse(r)[:, :, :n] for r in rhs]
    ex = [_modes_field(n, [((0, 1, 0), 0.0, -nu)]),
          _modes_field(n, [((0, 2, 0), 0.0, 1.0 / 3.0)]),
          np.zeros((n, n, n))]
    for g, e in zip(got, ex):
        assert np.abs(g - e).max() < 1e-13


def test_deviatoric_rhs_projects_to_divergence_form(oracle):
    """Band-limited fields: projected deviatoric RHS = projected divergence
    RHS to rounding, the difference is a gradient, and driver steps with
    the two forms give the same fields."""
    n, nu = 32, 0.02
    rng = np.random.default_rng(6)
    modes = [_random_modes(rng, [0, 1, 2, 3, -2], [0, 1, -1, 3], [0, 2, -1, -3], 6)
             for _ in range(3)]
    U = np.stack([oracle.fft_forward(_pad(_modes_field(n, md), n)) for md in modes])
    U = oracle.dealias_all(oracle.pressure(U, n), n // 3, n)
    r_div, _ = oracle.rhs_velocity(U, nu, 0, n)
    r_dev, _ = oracle.rhs_velocity_deviatoric(U, nu, n)
    pd, pv = oracle.pressure(r_div, n), oracle.pressure(r_dev, n)
    assert np.abs(pv - pd).max() < 1e-12 * np.abs(pd).max()
    d = [oracle.to_complex(r_dev[c] - r_div[c]) for c in range(3)]
    assert np.abs(d[0]).max() > 1e-3 * np.abs(pd).max()
    kx, ky, kz = oracle._kgrid(n)
    curl = [ky * d[2] - kz * d[1], kz * d[0] - kx * d[2], kx * d[1] - ky * d[0]]
    assert max(np.abs(c).max() for c in curl) < 1e-11 * np.abs(pd).max()
    out = {}
    for dev in (False, True):
        cfg = oracle.Config(n=n, nu=0.01, dt=0.002, dealias=0, flow_type=1, kfmax=2, famp=0.5,
                            deviatoric=dev)
        st = oracle.State(fields=U.copy(), rhs_old=np.zeros_like(U), dt=cfg.dt)
        for _ in range(3):
            st = oracle.step(cfg, st)
        out[dev] = st.fields
    assert np.abs(out[True] - out[False]).max() < 1e-12 * np.abs(out[False]).max()


# ------------------------------------------------------------ multi-step closed form

def _abc_field(oracle, n, a, b, c):
    """The ABC flow u = (a sin z + c cos y, b sin x + a cos z, c sin y + b cos x)
    as spectral fields.  It is a Beltrami field (curl u = u), so u x omega = 0
    and the nonlinear term -(u . grad) u = -grad(|u|^2/2) is a gradient that
    the projection removes: the flow is an exact Navier-Stokes solution whose
    modes (all |k| = 1) decay together as du/dt = -nu u."""
    comps = [[((0, 0, 1), 0.0, a), ((0, 1, 0), c, 0.0)],
             [((1, 0, 0), 0.0, b), ((0, 0, 1), a, 0.0)],
             [((0, 1, 0), 0.0, c), ((1, 0, 0), b, 0.0)]]
    return np.stack([oracle.fft_forward(_pad(_modes_field(n, md), n)) for md in comps])


def abc_amplitudes(steps, dt, rate):
    """Amplitude y_m of the field after m driver steps (y_0 = 1): an Euler
    first step, then AB2 (main.f90:124-176), for dy/dt = rate(y)."""
    ys, r_prev = [1.0], None
    for _ in range(steps):
        r = rate(ys[-1])
        ys.append(ys[-1] + dt * (r if r_prev is None else 1.5 * r - 0.5 * r_prev))
        r_prev = r
    return ys


@pytest.mark.parametrize("dealias,form,forced", [(0, "divergence", False), (1, "divergence", False),
                                                 (0, "rotational", False), (0, "deviatoric", False),
                                                 (0, "divergence", True)])
def test_step_abc_flow_closed_form(oracle, dealias, form, forced):
    """Twenty driver steps of the ABC flow, whose nonlinear term is nonzero
    (|grad(|u|^2/2)| ~ |u|^2) but a pure gradient: the fields stay f0 y_m,
    with y_m from the Euler/AB2 recurrence of dy/dt = -nu y (forced, KFMAX = 2:
    (FAMP/E - nu) y with the band energy E = E0 y^2, all of it in shell 1).
    A multi-step trajectory with a closed form: catches a nonlinear term left
    unprojected or projected before the update, an AB2 history that keeps
    the gradient part, and a forcing that drifts over steps."""
    n, nu, dt, steps, famp = 16, 0.05, 0.01, 20, 0.5
    a, b, c = 1.0, 0.8, 0.6
    f0 = _abc_field(oracle, n, a, b, c)
    e0 = (a * a + b * b + c * c) / 2.0
    cfg = oracle.Config(n=n, nu=nu, dt=dt, dealias=dealias, flow_type=1 if forced else 0,
                        kfmax=2, famp=famp, rotational=form == "rotational",
                        deviatoric=form == "deviatoric")
    rate = (lambda y: (famp / (e0 * y * y) - nu) * y) if forced else (lambda y: -nu * y)
    ys = abc_amplitudes(steps, dt, rate)
    nl, _ = oracle.rhs_velocity(f0, 0.0, 0, n)
    assert np.abs(nl).max() > 0.1 * np.abs(f0).max()   # the nonlinear term is not zero
    st = oracle.State(fields=f0.copy(), rhs_old=np.zeros_like(f0), dt=dt)
    scale = np.abs(f0).max()
    for m in range(1, steps + 1):
        st = oracle.step(cfg, st)
        assert np.abs(st.fields - ys[m] * f0).max() < 1e-12 * scale, m
        if forced:
            assert abs(st.energy - e0 * ys[m - 1] ** 2) < 1e-12 * e0
    assert abs(ys[-1] - 1.0) > 1e-3   # the amplitude has moved


def uniform_flow_scalar_case(oracle, n, u, q, sc, nu):
    """A uniform velocity u (the k = 0 mode, which neither the nonlinear term,
    the viscosity nor the projection changes) carrying two scalars
    phi = m + p cos(q.x) + r sin(q.x): type 0 (with the mean-gradient source
    -u_1, rhs_scalars.f90:387-389) and type 1.  Every entry k of a scalar's
    spectrum evolves as d(phi_k)/dt = lambda_k phi_k, lambda_k =
    -(i k.u + D |k|^2), D = nu/Sc, and the type-0 scalar's mean loses u_1
    per unit time.  Returns (fields, expected(m, dt) -> the scalars'
    spectra after m steps)."""
    # the velocity's k = 0 entries written directly (the transform of a
    # constant leaves roundoff in every other entry, which the flow then
    # carries: 4e-13 of |u| in physical space at 256^3)
    vel = np.zeros((3, n, n, n + 2))
    vel[:, 0, 0, 0] = np.asarray(u) * float(n) ** 3
    phys = [0.3 + _modes_field(n, [(q, 0.7, -0.4)]), -0.1 + _modes_field(n, [(q, 0.2, 0.5)])]
    fields = np.concatenate([vel, np.stack([oracle.fft_forward(_pad(p, n)) for p in phys])])
    kx, ky, kz = oracle._kgrid(n)
    k2 = (kx * kx + ky * ky) + kz * kz
    kdotu = (kx * u[0] + ky * u[1]) + kz * u[2]
    s0 = [oracle.to_complex(fields[3 + s]) for s in range(2)]

    def expected(m, dt):
        out = []
        for s in range(2):
            lam = -(1j * kdotu + (nu / sc[s]) * k2)
            a_prev, a = np.ones_like(lam), 1.0 + dt * lam   # Euler first step
            for _ in range(m - 1):
                a_prev, a = a, a + dt * (1.5 * lam * a - 0.5 * lam * a_prev)
            e = (a if m > 0 else np.ones_like(lam)) * s0[s]
            if s == 0:
                e[0, 0, 0] -= m * dt * u[0] * float(n) ** 3
            out.append(oracle.to_interleaved(e))
        return out

    return fields, expected


@pytest.mark.parametrize("dealias", [0, 1])
def test_step_uniform_flow_scalars_closed_form(oracle, dealias):
    """Twenty driver steps of two scalars in a uniform flow against the
    closed form of each spectral entry (Euler, then AB2, for the complex
    rate -(i k.u + D k^2)) and the mean-gradient source.  Catches the sign of
    the advection term, the diffusivity (nu/Sc), the source's component or
    sign, and the scalars' Euler/AB2 bookkeeping over many steps."""
    n, nu, dt, steps = 16, 0.02, 0.01, 20
    u, q, sc = (0.3, -0.2, 0.5), (2, 1, -3), (0.7, 2.0)
    f0, expected = uniform_flow_scalar_case(oracle, n, u, q, sc, nu)
    cfg = oracle.Config(n=n, nu=nu, dt=dt, dealias=dealias, flow_type=0, n_scalars=2, sc=sc,
                        scalar_type=(0, 1))
    st = oracle.State(fields=f0.copy(), rhs_old=np.zeros_like(f0), dt=dt)
    scale = np.abs(f0[3:]).max()
    for m in range(1, steps + 1):
        st = oracle.step(cfg, st)
        ex = expected(m, dt)
        for s in range(2):
            assert np.abs(st.fields[3 + s] - ex[s]).max() < 1e-12 * scale, (m, s)
        assert np.abs(st.fields[:3] - f0[:3]).max() < 1e-12 * np.abs(f0[:3]).max()
    # the scalars moved: advected (phase) and diffused (amplitude)
    assert np.abs(st.fields[3:] - f0[3:]).max() > 1e-2 * scale


def test_abc_spectra_written_directly_are_the_transforms(oracle):
    """The 1024^3 GPU check writes the ABC flow's half spectra entry by
    entry (`abc_spectral`); they must equal the forward transform of the
    physical fields, including the conjugate entries of the kx = 0 modes."""
    from test_gpu_abc import abc_physical, abc_spectral
    n = 16
    for p, s in zip(abc_physical(n), abc_spectral(n)):
        f = oracle.fft_forward(p)
        assert np.abs(f - s).max() < 1e-12 * np.abs(f).max()
