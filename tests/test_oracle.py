"""Pins the CPU oracle (oracle/hit3d_oracle.py, oracle/randu.c) against the
reference's own generator build, brute-force sums and closed forms.

Each test names the plausible mistake it would catch.
"""
import os
import subprocess

import numpy as np
import pytest

REF_RNG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "oracle", "_ref", "randu_ref")
RN1, RN2 = 67748937.0, 93712139.0


def _phys(n, f):
    dx = 2 * np.pi / n
    x = np.arange(n) * dx
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    out = np.zeros((n, n, n + 2))
    out[:, :, :n] = f(xx, yy, zz)
    return out


# --------------------------------------------------------------------------- RNG

@pytest.mark.skipif(not os.path.exists(REF_RNG), reason="reference RANDu.f build not present "
                    "(oracle/Makefile builds it where /root/reference exists)")
def test_rng_matches_reference_build(oracle):
    """The C restatement reproduces RANDu.f compiled with -i8 -r8 bit for bit,
    including the re-initialisation that keeps iy (init_scalars after
    init_velocity).  Catches a wrong constant, rounding of nint, or a reset iy."""
    out = subprocess.run([REF_RNG, repr(RN1), "200000", repr(RN2), "50000"], check=True,
                         capture_output=True, text=True).stdout.split()
    ref = np.array([int(v) for v in out], dtype=np.int64).view(np.float64)
    g = oracle.RanduGenerator()
    mine = [g.random(-RN1)]
    mine += list(g.draw(RN1, 200000))
    mine += [g.random(-RN2)]
    mine += list(g.draw(RN2, 50000))
    mine = np.array(mine)
    assert ref.shape == mine.shape
    assert np.array_equal(ref.view(np.int64), mine.view(np.int64))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "randu_golden.npz")


def test_rng_matches_golden_fixture(oracle):
    """The C restatement against committed draws of the reference's own
    RANDu.f build (tests/golden/make_randu_golden.py): random(-RN1), 3e6
    draws, then random(-RN2) on the same generator and 5e5 draws, sampled."""
    g = np.load(GOLDEN)
    rn1, rn2 = g["seeds"]
    c1, c2 = (int(v) for v in g["counts"])
    gen = oracle.RanduGenerator()
    run1 = np.concatenate([[gen.random(-rn1)], gen.draw(rn1, c1)])
    run2 = np.concatenate([[gen.random(-rn2)], gen.draw(rn2, c2)])
    assert np.array_equal(run1[g["idx1"]].view(np.int64), g["bits1"])
    assert np.array_equal(run2[g["idx2"]].view(np.int64), g["bits2"])


def test_rng_chain_matches_golden_fixture(oracle):
    """Every value of the call chain of a 16^3 run with two spectrum scalars
    (init_velocity on RN1, then random(-RN2) and its draws twice), from the
    reference's own RANDu.f build: pins that re-initialisation keeps iy."""
    g = np.load(os.path.join(os.path.dirname(GOLDEN), "randu_chain16.npz"))
    rn1, rn2 = g["seeds"]
    per = 18 * 16 * 16
    gen = oracle.RanduGenerator()
    vals = [gen.random(-rn1)] + list(gen.draw(rn1, 6 * per))
    for _ in range(2):
        vals += [gen.random(-rn2)] + list(gen.draw(rn2, 2 * per))
    assert np.array_equal(np.array(vals).view(np.int64), g["bits"])


def test_rng_known_values(oracle):
    """First values for the reference seed RN1 (as printed by the reference
    build: 0.65056227208012396 for random(-RN1))."""
    g = oracle.RanduGenerator()
    assert g.random(-RN1) == 0.65056227208012396
    d = g.draw(RN1, 2)
    assert d[0] == 0.91480268293909173
    assert d[1] == 0.33799488550497464


def test_box_muller_moments(oracle):
    """Box-Muller of the generator's draws is standard normal (catches a
    missing factor 2 or a cos/sin swap that changes nothing else)."""
    g = oracle.RanduGenerator()
    g.random(-RN1)
    u = g.draw(RN1, 400000)
    z = oracle.box_muller(u[:200000], u[200000:])
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    assert abs(np.mean(z ** 4) - 3.0) < 0.05


# ------------------------------------------------------------------- wavenumbers

def test_wavenumbers_n8(oracle):
    akx, aky, akz = oracle.wavenumbers(8)
    assert list(akx) == [0, 0, 1, 1, 2, 2, 3, 3, 4, 4]
    assert list(aky) == [0, 1, 2, 3, 4, -3, -2, -1]
    assert list(akz) == list(aky)


def test_ialias_n8_counts(oracle):
    """n=8: threshold real(8/3)=2; 3 of 8 ky, 3 of 8 kz, 4 of 10 i exceed it."""
    ia = oracle.ialias(8)
    assert ia.shape == (8, 8, 10)
    assert int(ia.sum()) == 3 * 8 * 10 + 3 * 8 * 10 + 8 * 8 * 4
    assert ia[0, 0, 0] == 0 and ia[3, 0, 0] == 1 and ia[3, 3, 6] == 3 and ia[0, 4, 9] == 2


def test_kmax(oracle):
    assert oracle.kmax_of(32, 0) == 10
    assert oracle.kmax_of(32, 1) == 15
    assert oracle.kmax_of(256, 1) == 120
    assert oracle.kmax_of(512, 0) == 170
    assert oracle.kmax_of(1024, 1) == 482


def test_shell_index_float32_equals_exact(oracle):
    """nint(sqrt(real(k2,4))) equals the exact integer rule for every k2 a
    1024^3 grid can produce (3*512^2); the HIP kernels use the integer rule."""
    k2 = np.arange(0, 3 * 512 ** 2 + 1)
    s = np.floor(np.sqrt(k2)).astype(np.int64)
    exact = np.where(k2 > s * s + s, s + 1, s)
    assert np.array_equal(oracle.shell_index_f32(k2), exact)
    assert np.array_equal(oracle.shell_index_f64(k2), exact)


# ---------------------------------------------------------------------------- FFT

def test_fft_single_mode_index_and_sign(oracle):
    """cos/sin of (2x + 3y - 5z) land at i=2*2 (re/im), j=kz index 11, k=ky 3
    with +n^3/2 and -n^3/2 (forward sign -1).  Catches swapped ky/kz or sign."""
    n = 16
    kx, ky, kz = 2, 3, -5
    c = oracle.fft_forward(_phys(n, lambda x, y, z: np.cos(kx * x + ky * y + kz * z)))
    s = oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(kx * x + ky * y + kz * z)))
    j, k = kz % n, ky % n
    assert abs(c[k, j, 2 * kx] - n ** 3 / 2) < 1e-9
    assert abs(s[k, j, 2 * kx + 1] + n ** 3 / 2) < 1e-9
    c[k, j, 2 * kx] = 0.0
    s[k, j, 2 * kx + 1] = 0.0
    assert np.abs(c).max() < 1e-9 and np.abs(s).max() < 1e-9


def test_fft_brute_force_dft(oracle):
    """Against a direct triple sum at n=6 (forward, and inverse with 1/n^3)."""
    n = 6
    rng = np.random.default_rng(1)
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = rng.standard_normal((n, n, n))
    spec = oracle.fft_forward(a)
    idx = np.arange(n)
    for ky in range(n):
        for kz in range(n):
            for kx in range(n // 2 + 1):
                ph = np.exp(-2j * np.pi * (kx * idx[None, None, :] + ky * idx[None, :, None]
                                           + kz * idx[:, None, None]) / n)
                v = np.sum(a[:, :, :n] * ph)
                assert abs(spec[ky, kz, 2 * kx] - v.real) < 1e-11
                assert abs(spec[ky, kz, 2 * kx + 1] - v.imag) < 1e-11
    back = oracle.fft_inverse(spec)
    assert np.abs(back[:, :, :n] - a[:, :, :n]).max() < 1e-13


def test_fft_inverse_ignores_imag_of_kx0_and_nyquist(oracle):
    """The c2r step uses only Re of the kx=0 and kx=n/2 terms after the y/z
    transforms (FFTW's c2r, x_fftw.f90:1066-1073): an imaginary part on a
    self-conjugate kx=0 mode must not change the output."""
    n = 8
    s = np.zeros((n, n, n + 2))
    s[0, 0, 1] = 5.0       # Im of the mean
    s[0, 0, n + 1] = 3.0   # Im of (kx=n/2, 0, 0)
    assert np.abs(oracle.fft_inverse(s)).max() == 0.0


def test_x_derivative(oracle):
    n = 16
    d = oracle.x_derivative(oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(2 * x) * np.cos(3 * z))), "z")
    back = oracle.fft_inverse(d)[:, :, :n]
    ex = _phys(n, lambda x, y, z: -3 * np.sin(2 * x) * np.sin(3 * z))[:, :, :n]
    assert np.abs(back - ex).max() < 1e-12


# ----------------------------------------------------------------------- RHS

def _sines(oracle, n):
    f = np.stack([oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(x))),
                  oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(y))),
                  oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(z)))])
    return f


@pytest.mark.parametrize("dealias", [0, 1])
@pytest.mark.parametrize("nu", [0.0, 0.3])
def test_rhs_velocity_analytic(oracle, dealias, nu):
    """test_rhs_velocity (rhs_velocity.f90:430-553): u=sin x, v=sin y, w=sin z
    gives RHS_u = -sin x (2 cos x + cos y + cos z + nu), etc.  Holds for both
    dealiasing branches (band-limited field).  Catches a wrong sign of the
    convective term, a missing factor 1/2 in the phase-shift average, or a
    transposed product index."""
    n = 16
    rhs, _ = oracle.rhs_velocity(_sines(oracle, n), nu, dealias, n)
    got = [oracle.fft_inverse(r)[:, :, :n] for r in rhs]
    ex = [_phys(n, lambda x, y, z: -np.sin(x) * (2 * np.cos(x) + np.cos(y) + np.cos(z) + nu)),
          _phys(n, lambda x, y, z: -np.sin(y) * (2 * np.cos(y) + np.cos(x) + np.cos(z) + nu)),
          _phys(n, lambda x, y, z: -np.sin(z) * (2 * np.cos(z) + np.cos(y) + np.cos(x) + nu))]
    for g, e in zip(got, ex):
        assert np.abs(g - e[:, :, :n]).max() < 1e-12


def test_rhs_velocity_asymmetric_field(oracle):
    """A field with different wavenumbers per component and a mixed product
    (u=sin(x+2y), v=cos(3z), w=sin(2x)) against its analytic -d_j(u_i u_j)."""
    n = 16
    U = np.stack([oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(x + 2 * y))),
                  oracle.fft_forward(_phys(n, lambda x, y, z: np.cos(3 * z))),
                  oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(2 * x)))])
    for dealias in (0, 1):
        rhs, _ = oracle.rhs_velocity(U, 0.0, dealias, n)
        got = [oracle.fft_inverse(r)[:, :, :n] for r in rhs]
        u = lambda x, y, z: np.sin(x + 2 * y)
        ux = lambda x, y, z: np.cos(x + 2 * y)
        uy = lambda x, y, z: 2 * np.cos(x + 2 * y)
        v = lambda x, y, z: np.cos(3 * z)
        vz = lambda x, y, z: -3 * np.sin(3 * z)
        w = lambda x, y, z: np.sin(2 * x)
        wx = lambda x, y, z: 2 * np.cos(2 * x)
        # -d_j(u_i u_j) = -(u_j d_j u_i + u_i div u); div u = ux
        ex = [lambda x, y, z: -(u(x, y, z) * ux(x, y, z) + v(x, y, z) * uy(x, y, z)
                                + u(x, y, z) * ux(x, y, z)),
              lambda x, y, z: -(w(x, y, z) * vz(x, y, z) + v(x, y, z) * ux(x, y, z)),
              lambda x, y, z: -(u(x, y, z) * wx(x, y, z) + w(x, y, z) * ux(x, y, z))]
        for gg, e in zip(got, ex):
            assert np.abs(gg - _phys(n, e)[:, :, :n]).max() < 1e-11


@pytest.mark.parametrize("dealias", [0, 1])
def test_rhs_scalars_analytic(oracle, dealias):
    """test_rhs_scalars (rhs_scalars.f90:504-594) with pe in place of nu:
    u=sin x, v=sin 5y, w=sin 17z, phi=cos x -> -cos 2x - cos x (5 cos 5y +
    17 cos 17z + pe).  Needs n >= 54 (mode 17 inside the 2/3 cube)."""
    n = 64
    pe = 0.5
    f = np.stack([oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(x))),
                  oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(5 * y))),
                  oracle.fft_forward(_phys(n, lambda x, y, z: np.sin(17 * z))),
                  oracle.fft_forward(_phys(n, lambda x, y, z: np.cos(x)))])
    r = oracle.rhs_scalars(f, 1, (pe,), (1,), dealias, n)
    got = oracle.fft_inverse(r[0])[:, :, :n]
    ex = _phys(n, lambda x, y, z: -np.cos(2 * x) - np.cos(x) * (5 * np.cos(5 * y)
                                                                 + 17 * np.cos(17 * z) + pe))
    assert np.abs(got - ex[:, :, :n]).max() < 1e-10
    # scalar type 0 adds the mean-gradient source -u (rhs_scalars.f90:387-389)
    r0 = oracle.rhs_scalars(f, 1, (pe,), (0,), dealias, n)
    src = oracle.fft_inverse(r0[0] - r[0])[:, :, :n]
    assert np.abs(src + np.sin(_phys(n, lambda x, y, z: x)[:, :, :n])).max() < 1e-12


# -------------------------------------------------------- projection, forcing

def _random_spec(oracle, n, seed, count=3):
    rng = np.random.default_rng(seed)
    out = np.zeros((count, n, n, n + 2))
    for c in range(count):
        a = np.zeros((n, n, n + 2))
        a[:, :, :n] = rng.standard_normal((n, n, n))
        out[c] = oracle.fft_forward(a)
    return out


def _spec_at(c, kx):
    """complex entry (kx, kz=0, ky=0) of an interleaved spectral slot."""
    return c[0, 0, 2 * kx] + 1j * c[0, 0, 2 * kx + 1]


@pytest.mark.parametrize("st", [100, 200, 300])
def test_scalar_reaction_closed_forms(oracle, st):
    """Zero velocity, Pe = 0, phi = a cos 3x on n = 16 (kmax 5): the RHS is the
    transformed reaction rate.  KPP r(1 - phi^2) = r(1 - a^2/2) - (r a^2/2) cos 6x
    (cos 6x is outside the cube and is cut by dealias_rhs); the bistable forms
    r(1-phi^2)phi give (r a - 3/4 r a^3) cos 3x, and their cos 9x part aliases
    to kx = 7 on this grid, which dealias_rhs also removes."""
    n, a, r = 16, 0.6, 2.5
    n3 = float(n) ** 3
    z = np.zeros((n, n, n + 2))
    phi = oracle.fft_forward(_phys(n, lambda x, y, zz: a * np.cos(3 * x)))
    f = np.stack([z, z, z, phi])
    out = oracle.rhs_scalars(f, 1, [0.0], [st], 0, n, reac=[r])[0]
    c = oracle.to_complex(out)
    if st == 100:
        want0, want3 = r * (1 - a * a / 2), 0.0
    else:
        want0, want3 = 0.0, (r * a - 0.75 * r * a ** 3) / 2
    assert abs(_spec_at(out, 0) - n3 * want0) < 1e-9
    assert abs(_spec_at(out, 3) - n3 * want3) < 1e-9
    c[0, 0, 0] = 0.0
    c[0, 0, 3] = 0.0
    assert np.abs(c).max() < 1e-9


def test_scalar_reaction_311_phase_shift(oracle):
    """Type 311 with the phase shift: the scalar is cut at |k_i| > n/4 before
    the rate r(1-phi^2)(phi-mean) is formed, and nothing is cut after it:
    a cos 3x gives (r a - 3/4 r a^3)/2 at kx = 3 and the aliased -r a^3/8 at
    kx = 7; a cos 5x is cut entirely, so the RHS vanishes."""
    n, a, r = 16, 0.6, 2.5
    n3 = float(n) ** 3
    z = np.zeros((n, n, n + 2))
    phi = oracle.fft_forward(_phys(n, lambda x, y, zz: a * np.cos(3 * x)))
    out = oracle.rhs_scalars(np.stack([z, z, z, phi]), 1, [0.0], [311], 1, n, reac=[r])[0]
    assert abs(_spec_at(out, 3) - n3 * (r * a - 0.75 * r * a ** 3) / 2) < 1e-9
    assert abs(_spec_at(out, 7) - n3 * (-r * a ** 3 / 8)) < 1e-9
    phi5 = oracle.fft_forward(_phys(n, lambda x, y, zz: a * np.cos(5 * x)))
    out = oracle.rhs_scalars(np.stack([z, z, z, phi5]), 1, [0.0], [311], 1, n, reac=[r])[0]
    assert np.abs(out).max() < 1e-9


def test_pressure_projection(oracle):
    """k.u = 0 after projection, projection is idempotent and removes a pure
    gradient (u = grad phi) entirely.  Catches k_y/k_z swaps and sign errors."""
    n = 16
    U = oracle.pressure(_random_spec(oracle, n, 2), n)
    kx, ky, kz = oracle._kgrid(n)
    C = [oracle.to_complex(U[c]) for c in range(3)]
    div = kx * C[0] + ky * C[1] + kz * C[2]
    assert np.abs(div).max() < 1e-9 * np.abs(C[0]).max()
    assert np.abs(oracle.pressure(U, n) - U).max() < 1e-10 * np.abs(U).max()
    phi = oracle.to_complex(_random_spec(oracle, n, 3, 1)[0])
    grad = np.stack([oracle.to_interleaved(1j * k * phi) for k in (kx, ky + 0 * kx, kz + 0 * kx)])
    assert np.abs(oracle.pressure(grad, n)).max() < 1e-9 * np.abs(grad).max()


def test_force_single_shell(oracle):
    """u = cos x has band energy 1/4 in shell 1 (Parseval with the reference's
    1/n^6 normalisation); forcing adds famp/E * u-hat on that mode only."""
    n = 16
    f = np.stack([oracle.fft_forward(_phys(n, lambda x, y, z: np.cos(x))),
                  np.zeros((n, n, n + 2)), np.zeros((n, n, n + 2))])
    rhs, energy = oracle.force_velocity(f, np.zeros_like(f), 2, 0.5, n)
    assert abs(energy - 0.25) < 1e-14
    assert np.abs(rhs - (0.5 / 0.25) * f).max() < 1e-9


def test_forced_node_count(oracle):
    """KFMAX=2 forces the 50 complex modes with 0 < |k| < 2.5, kx >= 0 half
    space counted as in m_force.f90:37-46 (re and im entries -> 100)."""
    n = 32
    m = oracle.forced_mask(n, 2)
    assert int(m.sum()) == 2 * 50


def test_stats_single_modes(oracle):
    """E(k) of cos y (kx=0 entries halved) and cos x are both 1/4 in shell 1;
    w = sin x has enstrophy <|curl u|^2> = 1/2."""
    n = 16
    z = np.zeros((n, n, n + 2))
    for fun in (lambda x, y, zz: np.cos(y), lambda x, y, zz: np.cos(x)):
        f = np.stack([oracle.fft_forward(_phys(n, fun)), z, z])
        st = oracle.stats(f, n, 5, 0.01, 100.0)
        assert abs(st["e_spec"][0] - 0.25) < 1e-14 and abs(st["energy"] - 0.25) < 1e-14
    f = np.stack([z, z, oracle.fft_forward(_phys(n, lambda x, y, zz: np.sin(x)))])
    st = oracle.stats(f, n, 5, 0.01, 100.0)
    assert abs(st["enstrophy"] - 0.5) < 1e-13
    # dissipation 2 nu k^2 E(k)
    assert abs(st["eps_v"] - 2 * 0.01 * 0.25) < 1e-15


def test_shell_spectrum_chunked_equals_whole(oracle):
    """The plane-chunked shell sums (used on 1024^3 fields) equal the
    whole-array ones to rounding, for chunks that do and do not divide n."""
    n = 16
    rng = np.random.default_rng(9)
    f = rng.standard_normal((3, n, n, n + 2))
    whole = oracle.shell_spectrum(f, n, 12)
    for planes in (1, 5, 16):
        assert np.allclose(oracle.shell_spectrum_chunked(f, n, 12, planes), whole,
                           rtol=1e-13, atol=0)


def test_stats_parseval(oracle):
    """Sum of E(k) over all shells equals 1/2 <|u|^2> for a field truncated to
    a sphere inside the shells counted."""
    n = 16
    U = _random_spec(oracle, n, 4)
    k2 = oracle._k2_interleaved(n)
    sh = oracle.shell_index_f32(k2)
    U = np.where(((sh <= 6) & (sh > 0))[None] & (np.arange(n + 2) < n)[None, None, None, :], U, 0.0)
    e = oracle.shell_spectrum(U, n, 6)
    phys = [oracle.fft_inverse(U[c])[:, :, :n] for c in range(3)]
    half = 0.5 * np.mean(phys[0] ** 2 + phys[1] ** 2 + phys[2] ** 2)
    assert abs(e[1:].sum() + e[0] - half) < 1e-12 * half


def test_scalar_stats_closed_form(oracle):
    """phi = c + a cos x + b sin 2y: mean c, variance (a^2+b^2)/2, E(1) = a^2/2,
    E(2) = b^2/2 (catches a missing doubling or kx=0 halving), dissipation
    2 pe <|grad phi|^2> = 2 pe (a^2/2 + 2 b^2), min/max c -+ (a+b) on grid
    points x = pi, y = 3pi/4."""
    n, a, b, c, pe = 16, 0.7, 0.3, 1.25, 0.02
    phi = oracle.fft_forward(_phys(n, lambda x, y, z: c + a * np.cos(x) + b * np.sin(2 * y)))
    st = oracle.scalar_stats(phi, n, 5, pe)
    assert abs(st["mean"] - c) < 1e-14
    assert abs(st["variance"] - (a * a + b * b) / 2) < 1e-13
    es = st["es_spec"]
    assert es.shape == (5,)
    assert abs(es[0] - a * a / 2) < 1e-14 and abs(es[1] - b * b / 2) < 1e-14
    assert np.abs(es[2:]).max() < 1e-28
    assert abs(st["sc_diss"] - 2 * pe * (a * a / 2 + 2 * b * b)) < 1e-14
    assert abs(st["sc_min"] - (c - a - b)) < 1e-14 and abs(st["sc_max"] - (c + a + b)) < 1e-14


def test_scalar_spectrum_sums_to_variance(oracle):
    """For a field truncated inside the counted shells, sum E(k) = variance
    (Parseval with the half-spectrum weights)."""
    n = 16
    phi = _random_spec(oracle, n, 7, count=1)[0]
    sh = oracle.shell_index_f32(oracle._k2_interleaved(n))
    phi = np.where((sh <= 5) & (np.arange(n + 2) < n)[None, None, :], phi, 0.0)
    st = oracle.scalar_stats(phi, n, 5, 0.1)
    assert abs(st["es_spec"].sum() - st["variance"]) < 1e-12 * st["variance"]


def test_skewness_closed_form(oracle):
    """u = (sin x + a sin 2x, 0, 0)-type fields: du/dx = cos x + 2a cos 2x has
    <d^2> = 1/2 + 2a^2 and <d^3> = 3a/2 (only the cos^2 x cos 2x term
    survives), the other components have zero derivative variance and are
    given a symmetric profile (zero skewness)."""
    n, a = 32, 0.3
    f = np.stack([oracle.fft_forward(_phys(n, g)) for g in (
        lambda x, y, z: np.sin(x) + a * np.sin(2 * x),
        lambda x, y, z: np.sin(y), lambda x, y, z: np.sin(z))])
    sk = oracle.skewness(f, n)
    want = ((1.5 * a) / (0.5 + 2 * a * a) ** 1.5 + 0.0 + 0.0) / 3.0
    assert abs(sk - want) < 1e-13


def test_divergence_closed_form(oracle):
    """u = (a sin x, b sin 2y, c sin 4z): div u = a cos x + 2b cos 2y + 4c cos 4z,
    extremes -+(a + 2b + 4c) at grid points; zero after projection."""
    n, a, b, c = 16, 0.5, -0.25, 0.125
    f = np.stack([oracle.fft_forward(_phys(n, g)) for g in (
        lambda x, y, z: a * np.sin(x), lambda x, y, z: b * np.sin(2 * y),
        lambda x, y, z: c * np.sin(4 * z))])
    lo, hi = oracle.divergence(f, n)
    m = abs(a) + 2 * abs(b) + 4 * abs(c)
    assert abs(lo + m) < 1e-14 and abs(hi - m) < 1e-14
    lo, hi = oracle.divergence(oracle.pressure(_random_spec(oracle, n, 3), n), n)
    assert max(abs(lo), abs(hi)) < 1e-12


# ------------------------------------------------------------- scalar init

def test_init_scalars_spectrum_shapes(oracle):
    """Spectrum scalars come out with the normalised target spectrum: the
    scalar spectrum of stat_scalars (2 x shell sums) equals 2 x target in every
    shell (2/3 rule, where the ialias mask leaves shells <= kmax intact)."""
    n = 32
    kmax = oracle.kmax_of(n, 0)
    gen = oracle.RanduGenerator()
    oracle.init_velocity_uniforms(n, RN1, gen)
    f = oracle.init_scalars(n, kmax, [1, 2], [3.0, 2.0], RN2, gen)
    for s, st in enumerate((1, 2)):
        es = oracle.scalar_stats(f[s], n, kmax, 1.0)["es_spec"]
        tgt = oracle.scalar_target_spectrum(kmax, st, [3.0, 2.0][s])
        assert np.abs(es - 2 * tgt).max() < 1e-12 * tgt.max()
        assert abs(tgt.sum() - 1.0) < 1e-14


def test_init_scalars_space_types(oracle):
    """Type 13 is exactly sin(k0 x) with k0 = min(peak, n/8); type 11 varies in z
    only and type 12 in x only, both with zero mean and values in [-1, 1]."""
    n = 32
    kmax = oracle.kmax_of(n, 0)
    f = oracle.init_scalars(n, kmax, [13, 11, 12], [7.0, 2.0, 2.0], RN2, oracle.RanduGenerator())
    x = np.arange(n) * 2 * np.pi / n
    p13 = oracle.fft_inverse(f[0])[:, :, :n]
    assert np.abs(p13 - np.sin(4.0 * x)[None, None, :]).max() < 1e-14
    c11 = oracle.to_complex(f[1])
    nz = np.argwhere(np.abs(c11) > 1e-9 * np.abs(c11).max())
    assert (nz[:, 0] == 0).all() and (nz[:, 2] == 0).all()        # ky = kx = 0: z only
    c12 = oracle.to_complex(f[2])
    nz = np.argwhere(np.abs(c12) > 1e-9 * np.abs(c12).max())
    assert (nz[:, 0] == 0).all() and (nz[:, 1] == 0).all()        # ky = kz = 0: x only
    for c in (c11, c12):
        assert c[0, 0, 0] == 0.0
    p11 = oracle.fft_inverse(f[1])[:, :, :n]
    assert np.abs(p11).max() < 1.2


def test_init_scalars_double_delta(oracle):
    """Type 3: shells 0 and >= 2 kmax/3 are empty after smoothing, and the field
    is far from the Gaussian it started as: its flatness <phi^4>/<phi^2>^2 is
    near 2 (a smoothed two-valued field), where type 1 gives 3."""
    n = 32
    kmax = oracle.kmax_of(n, 0)
    gen = oracle.RanduGenerator()
    f = oracle.init_scalars(n, kmax, [3], [3.0], RN2, gen)[0]
    sh = oracle.shell_index_f32(oracle._k2_interleaved(n))
    assert np.abs(np.where((sh == 0) | (sh >= kmax * 2 // 3), f, 0.0)).max() == 0.0
    p = oracle.fft_inverse(f)[:, :, :n]
    assert np.mean(p ** 4) / np.mean(p ** 2) ** 2 < 2.4


# ------------------------------------------------------------------ particles

def _index_field(n, fun):
    """physical slot with a[z, y, x] = fun(x+1, y+1, z+1) (1-based grid units)."""
    i = np.arange(1, n + 1, dtype=np.float64)
    zz, yy, xx = np.meshgrid(i, i, i, indexing="ij")
    out = np.zeros((n, n, n + 2))
    out[:, :, :n] = fun(xx, yy, zz)
    return out


def test_particles_trilinear_reproduces_linear(oracle):
    """Trilinear weights reproduce a + bx + cy + dz + e xyz-free bilinear terms
    exactly inside one cell (stencil away from the periodic seam)."""
    n = 16
    fun = lambda x, y, z: 0.3 + 0.25 * x - 0.5 * y + 0.125 * z + 0.0625 * x * y
    vel = [_index_field(n, fun), _index_field(n, lambda x, y, z: 2.0 * x),
           _index_field(n, lambda x, y, z: -z)]
    rng = np.random.default_rng(5)
    xyz = rng.uniform(3.0, 12.0, (200, 3))
    new, uvw = oracle.particles_move(vel, xyz, np.full((200, 3), -999.0), 0.0, n, 0)
    x, y, z = xyz.T
    assert np.abs(uvw[:, 0] - fun(x, y, z)).max() < 1e-13
    assert np.abs(uvw[:, 1] - 2 * x).max() < 1e-13 and np.abs(uvw[:, 2] + z).max() < 1e-13
    assert np.array_equal(new, xyz)


def test_particles_cint_reproduces_cubic(oracle):
    """The CINT stencil (nodes -1..2 around the cell) reproduces products of
    cubics exactly; a shifted or mis-weighted node breaks this."""
    n = 16
    px = lambda x: 0.5 - 0.1 * x + 0.02 * x ** 2 - 0.003 * x ** 3
    py = lambda y: 1.0 + 0.05 * y ** 3
    pz = lambda z: 0.2 * z - 0.01 * z ** 2
    fun = lambda x, y, z: px(x) * py(y) * pz(z)
    vel = [_index_field(n, fun)] * 3
    rng = np.random.default_rng(6)
    xyz = rng.uniform(3.0, 13.0, (300, 3))
    _, uvw = oracle.particles_move(vel, xyz, np.full((300, 3), -999.0), 0.0, n, 1)
    want = fun(*xyz.T)
    assert np.abs(uvw[:, 0] - want).max() < 1e-12 * np.abs(want).max()
    # trilinear does not reproduce the cubic (the check has teeth)
    _, uvw0 = oracle.particles_move(vel, xyz, np.full((300, 3), -999.0), 0.0, n, 0)
    assert np.abs(uvw0[:, 0] - want).max() > 1e-4 * np.abs(want).max()


@pytest.mark.parametrize("scheme", [0, 1])
def test_particles_uniform_flow_euler_ab2_and_wrap(oracle, scheme):
    """Uniform velocity: each move advances by dt*U/dx (Euler first, then AB2
    with 1.5U - 0.5U = U), wrapping periodically in x, y and z."""
    n = 16
    dx = 2 * np.pi / n
    U = (0.7, -0.4, 0.9)
    vel = [_index_field(n, lambda x, y, z, c=c: np.full_like(x, c)) for c in U]
    xyz = np.array([[16.9, 1.05, 16.95], [1.2, 8.0, 1.01], [8.5, 8.5, 8.5]])
    uvw = np.full((3, 3), -999.0)
    dt = 0.6 * dx / 0.9 * 0.3
    pos = xyz.copy()
    for step in range(6):
        pos, uvw = oracle.particles_move(vel, pos, uvw, dt, n, scheme)
        assert np.allclose(uvw, np.array(U)[None, :], rtol=0, atol=1e-14)
        want = (xyz - 1.0 + (step + 1) * dt * np.array(U) / dx) % n + 1.0
        assert np.abs(pos - want).max() < 1e-12
        assert (pos >= 1.0).all() and (pos < n + 1.0).all()


def test_particles_lattice(oracle):
    """nptot rounds to the nearest cube (m_particles.f90:222-228); lattice at
    cell centres of a 2 pi / npthird spacing, in grid units."""
    n = 32
    nptot, xyz = oracle.particles_lattice(1100, n)
    assert nptot == 1000 and xyz.shape == (1000, 3)
    assert oracle.particles_lattice(1200, n)[0] == 1331
    assert np.allclose(xyz[0], 0.5 * n / 10 + 1.0) and np.allclose(xyz[1, 0], 1.5 * n / 10 + 1.0)
    assert np.allclose(xyz[10, 1], 1.5 * n / 10 + 1.0) and np.allclose(xyz[100, 2], 1.5 * n / 10 + 1.0)


# ------------------------------------------------------------------ update, dt

def test_advance_euler_then_ab2(oracle):
    """With a constant RHS, Euler then AB2 advances by exactly dt*r each step
    (1.5 r - 0.5 r = r); catches swapped AB2 weights."""
    rng = np.random.default_rng(5)
    f0 = rng.standard_normal((3, 4, 4, 6))
    r = rng.standard_normal((3, 4, 4, 6))
    f1, ro = oracle.advance(f0, r, np.zeros_like(r), 0.1, True)
    f2, ro = oracle.advance(f1, r, ro, 0.1, False)
    assert np.allclose(f2, f0 + 0.2 * r, rtol=0, atol=1e-14)
    f3, _ = oracle.advance(f2, 2 * r, ro, 0.1, False)
    assert np.allclose(f3, f2 + 0.1 * (3.0 * r - 0.5 * r), rtol=0, atol=1e-14)


def test_my_dt(oracle):
    assert oracle.my_dt(0.001, 0.05) == 0.001 * 1.01
    assert oracle.my_dt(0.0099999, 0.05) == 0.01
    assert oracle.my_dt(0.002, 0.2) == 0.002 / 1.05
    assert oracle.my_dt(0.002, 0.05, True, (2.0,)) == min(0.002 * 1.01, 0.002 * 1.01 * 0.1 / 0.05 / 2.0)


# --------------------------------------------------------------------- init

def test_init_velocity_properties(oracle):
    """The seeded field is solenoidal, zero outside shells 1..kmax and has
    exactly the target shell spectrum (init_velocity.f90:123-237)."""
    n = 16
    kmax = oracle.kmax_of(n, 0)
    f, g, e0 = oracle.init_velocity(n, RN1, kmax, 2, 3.0, 4.0)
    kx, ky, kz = oracle._kgrid(n)
    C = [oracle.to_complex(f[c]) for c in range(3)]
    assert np.abs(kx * C[0] + ky * C[1] + kz * C[2]).max() < 1e-9 * np.abs(C[0]).max()
    e = oracle.shell_spectrum(f, n, kmax)[1:]
    tgt = oracle.target_spectrum(kmax, 2, 3.0, 4.0)
    assert np.allclose(e, tgt, rtol=1e-12, atol=0)
    # target type 2 is k^4 exp(-2 (k/3)^2), normalised
    k = np.arange(1, kmax + 1)
    t2 = k ** 4.0 * np.exp(-2.0 * (k / 3.0) ** 2)
    assert np.allclose(tgt, t2 / t2.sum(), rtol=1e-13)
    # the Gaussian field is what Box-Muller makes of the first six slots
    uni, _ = oracle.init_velocity_uniforms(n, RN1)
    assert np.array_equal(g, oracle.box_muller(uni[0:3], uni[3:6]))


# ------------------------------------------------------------------- Smagorinsky

def _shear_field(oracle, n):
    """u = (sin y, 0, cos y): divergence-free with |S| = 1 everywhere, so the
    Smagorinsky viscosity is the constant (c_s dx)^2."""
    z = np.zeros((n, n, n + 2))
    return np.stack([oracle.fft_forward(_phys(n, lambda x, y, zz: np.sin(y))), z,
                     oracle.fft_forward(_phys(n, lambda x, y, zz: np.cos(y)))])


def test_smagorinsky_viscosity_closed_form(oracle):
    n = 16
    tv = oracle.turb_visc_smag(_shear_field(oracle, n), n)
    dx = 2 * np.pi / n
    assert np.abs(tv - (0.18 * dx) ** 2).max() < 1e-15


def test_smagorinsky_velocity_source_closed_form(oracle):
    """With constant nu_t the stress divergence is nu_t lap(u) = -nu_t u for
    this field: catches a sign error or a missing factor 2 in tau = -2 nu_t S."""
    n = 16
    f = _shear_field(oracle, n)
    tv = oracle.turb_visc_smag(f, n)
    r = oracle.les_rhs_velocity(f, np.zeros_like(f), tv, n)
    nut = (0.18 * 2 * np.pi / n) ** 2
    for c in range(3):
        assert np.abs(r[c] + nut * f[c]).max() < 1e-12 * n ** 3


def test_smagorinsky_scalar_source_closed_form(oracle):
    n = 16
    tv = np.full((n, n, n), 0.3)
    phi = oracle.fft_forward(_phys(n, lambda x, y, z: np.cos(x) + np.sin(2 * z)))
    r = oracle.les_rhs_scalar(phi, np.zeros_like(phi), tv, n)
    lap = oracle.fft_forward(_phys(n, lambda x, y, z: -np.cos(x) - 4 * np.sin(2 * z)))
    assert np.abs(r - 0.3 * lap).max() < 1e-12 * n ** 3
