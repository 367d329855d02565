"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on
the same seeded inputs.

Tolerances (FP64 path): transforms 1e-13 relative (max-abs error over
max-abs value), one right-hand side 1e-12 relative, masks, wavenumber
indexing and the seeded Gaussian field bit-exact.
"""
import sys

import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS, RHS_OLD, WRK

pytestmark = pytest.mark.gpu

RN1 = 67748937.0


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def phys_random(n, seed, count=1):
    rng = np.random.default_rng(seed)
    out = np.zeros((count, n, n, n + 2))
    out[:, :, :, :n] = rng.uniform(-1, 1, (count, n, n, n))
    return out


def solenoidal_fields(oracle, n, seed, dealias, count=3):
    """Random dealiased, projected velocity (and extra scalar slots)."""
    ph = phys_random(n, seed, count)
    f = np.stack([oracle.fft_forward(p) for p in ph])
    f[:3] = oracle.pressure(f[:3], n)
    kmax = oracle.kmax_of(n, dealias)
    f = oracle.dealias_all(f, kmax, n)
    if dealias == 1:  # phase shift keeps modes with at most one large component
        ia = oracle.ialias(n)
        f = np.where(ia[None] >= 2, 0.0, f)
    return f


# ------------------------------------------------------------------ transforms

@pytest.mark.parametrize("n", [8, 16, 32, 64, 128])
def test_fft_matches_oracle(oracle, n):
    note(f"fft n={n}")
    h = hit3d_amd.Hit3D(n)
    a = phys_random(n, 10 + n)[0]
    h.put(WRK, 2, a)
    h.xFFT3d(1, 2)
    spec = h.get(WRK, 2)
    ref = oracle.fft_forward(a)
    assert rel(spec, ref) < 1e-13
    h.xFFT3d(-1, 2)
    back = h.get(WRK, 2)
    assert rel(back[:, :, :n], a[:, :, :n]) < 1e-14
    # inverse of an arbitrary spectral slot (non-Hermitian kx=0 plane included)
    s = phys_random(n, 20 + n)[0] * n ** 1.5
    h.put(WRK, 3, s)
    h.xFFT3d(-1, 3)
    assert rel(h.get(WRK, 3)[:, :, :n], oracle.fft_inverse(s)[:, :, :n]) < 1e-13
    h.close()


def test_fft_fields_and_slot_bookkeeping(oracle):
    """xFFT3d_fields and xFFT3d on different slots do not disturb each other
    (the transforms swap device buffers internally)."""
    n = 32
    note("fft fields n=32")
    h = hit3d_amd.Hit3D(n, n_scalars=1)
    a = phys_random(n, 3, 4)
    for c in range(4):
        h.put(FIELDS, c + 1, a[c])
    for c in range(7):
        h.put(WRK, c, a[c % 4])
    for c in range(1, 5):
        h.xFFT3d_fields(1, c)
    h.xFFT3d(1, 0)
    h.xFFT3d(1, 5)
    for c in range(4):
        assert rel(h.get(FIELDS, c + 1), oracle.fft_forward(a[c])) < 1e-13
    assert rel(h.get(WRK, 0), oracle.fft_forward(a[0])) < 1e-13
    assert rel(h.get(WRK, 5), oracle.fft_forward(a[1])) < 1e-13
    assert np.array_equal(h.get(WRK, 1)[:, :, :n], a[1][:, :, :n])
    h.close()


def test_fft_single_mode_device(oracle):
    n = 32
    note("single mode n=32")
    h = hit3d_amd.Hit3D(n)
    dx = 2 * np.pi / n
    x = np.arange(n) * dx
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = np.sin(5 * xx - 7 * yy + 11 * zz)
    h.put(WRK, 1, a)
    h.xFFT3d(1, 1)
    s = h.get(WRK, 1)
    k, j = (-7) % n, 11
    assert abs(s[k, j, 11] + n ** 3 / 2) < 1e-9 * n ** 3
    s[k, j, 11] = 0.0
    assert np.abs(s).max() < 1e-11 * n ** 3
    h.close()


@pytest.mark.parametrize("n", [256])
def test_fft_large_roundtrip(oracle, n):
    note(f"fft large n={n}")
    h = hit3d_amd.Hit3D(n)
    a = phys_random(n, 7)[0]
    h.put(WRK, 1, a)
    h.xFFT3d(1, 1)
    assert rel(h.get(WRK, 1), oracle.fft_forward(a)) < 1e-13
    h.xFFT3d(-1, 1)
    assert rel(h.get(WRK, 1)[:, :, :n], a[:, :, :n]) < 1e-14
    h.close()


def test_x_derivative(oracle):
    n = 32
    note("x_derivative")
    h = hit3d_amd.Hit3D(n)
    s = oracle.fft_forward(phys_random(n, 5)[0])
    h.put(WRK, 1, s)
    for ax, nto in (("x", 2), ("y", 3), ("z", 4)):
        h.x_derivative(1, ax, nto)
        assert np.array_equal(h.get(WRK, nto)[:, :, :n + 2], oracle.x_derivative(s, ax))
    with pytest.raises(hit3d_amd.Hit3dError):
        h.x_derivative(1, "q", 2)
    h.close()


@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_ialias_bit_exact(oracle, n):
    note(f"ialias n={n}")
    h = hit3d_amd.Hit3D(n)
    assert np.array_equal(h.ialias(), oracle.ialias(n))
    h.close()


@pytest.mark.parametrize("dealias", [0, 1])
def test_dealias_all(oracle, dealias):
    n = 32
    note(f"dealias_all {dealias}")
    h = hit3d_amd.Hit3D(n, n_scalars=1, dealias=dealias)
    f = np.stack([oracle.fft_forward(p) for p in phys_random(n, 9, 4)])
    for c in range(4):
        h.put(FIELDS, c + 1, f[c])
    h.dealias_all()
    ref = oracle.dealias_all(f, oracle.kmax_of(n, dealias), n)
    for c in range(4):
        assert np.array_equal(h.get(FIELDS, c + 1), ref[c])
    h.close()


# ------------------------------------------------------------------ physics

def test_pressure(oracle):
    n = 32
    note("pressure")
    h = hit3d_amd.Hit3D(n)
    f = np.stack([oracle.fft_forward(p) for p in phys_random(n, 11, 3)])
    for c in range(3):
        h.put(FIELDS, c + 1, f[c])
    h.pressure()
    ref = oracle.pressure(f, n)
    for c in range(3):
        assert rel(h.get(FIELDS, c + 1), ref[c]) < 1e-15
    h.close()


def test_advance(oracle):
    n = 16
    note("advance")
    h = hit3d_amd.Hit3D(n, n_scalars=1)
    rng = np.random.default_rng(4)
    f = rng.standard_normal((4, n, n, n + 2))
    r = rng.standard_normal((4, n, n, n + 2))
    for c in range(4):
        h.put(FIELDS, c + 1, f[c])
        h.put(WRK, c + 1, r[c])
    first = h.advance(1, True, 0.01)
    assert first is False
    fr, ro = oracle.advance(f[:3], r[:3], None, 0.01, True)
    for c in range(3):  # the device may contract f + dt*r into one fma
        assert rel(h.get(FIELDS, c + 1), fr[c]) < 1e-15
        assert np.array_equal(h.get(RHS_OLD, c + 1), ro[c])
    r2 = rng.standard_normal((4, n, n, n + 2))
    for c in range(4):
        h.put(WRK, c + 1, r2[c])
    h.advance(1, False, 0.01)
    fr2, _ = oracle.advance(fr, r2[:3], ro, 0.01, False)
    for c in range(3):
        assert rel(h.get(FIELDS, c + 1), fr2[c]) < 1e-15
    h.advance(2, True, 0.02)
    assert rel(h.get(FIELDS, 4), f[3] + 0.02 * r2[3]) < 1e-15
    h.close()


@pytest.mark.parametrize("dealias", [0, 1])
@pytest.mark.parametrize("n", [16, 32])
def test_rhs_velocity(oracle, n, dealias):
    note(f"rhs_velocity n={n} dealias={dealias}")
    h = hit3d_amd.Hit3D(n, dealias=dealias)
    nu = 0.01
    h.set_physics(nu)
    f = solenoidal_fields(oracle, n, 30 + n, dealias)
    for c in range(3):
        h.put(FIELDS, c + 1, f[c])
    h.rhs_scalars(False)
    phys_ref = [oracle.fft_inverse(f[c]) for c in range(3)]
    for c in range(3):
        assert rel(h.get(WRK, c + 1)[:, :, :n], phys_ref[c][:, :, :n]) < 1e-13
    dt = 0.005
    courant = h.rhs_velocity(dt)
    rhs, phys = oracle.rhs_velocity(f, nu, dealias, n)
    assert abs(courant - oracle.courant_number(phys, n, dt)) < 1e-13 * courant
    for c in range(3):
        assert rel(h.get(WRK, c + 1), rhs[c]) < 1e-12
    h.close()


@pytest.mark.parametrize("dealias", [0, 1])
def test_rhs_scalars(oracle, dealias):
    n = 32
    note(f"rhs_scalars dealias={dealias}")
    h = hit3d_amd.Hit3D(n, n_scalars=2, dealias=dealias)
    nu = 0.02
    sc = (0.7, 1.0)
    pe = tuple(nu / s for s in sc)
    h.set_physics(nu, pe=pe, scalar_type=(0, 1))
    f = solenoidal_fields(oracle, n, 40, dealias, count=5)
    for c in range(5):
        h.put(FIELDS, c + 1, f[c])
    h.rhs_scalars(True)
    ref = oracle.rhs_scalars(f, 2, pe, (0, 1), dealias, n)
    for s in range(2):
        assert rel(h.get(WRK, 4 + s), ref[s]) < 1e-12
    for c in range(3):
        assert rel(h.get(WRK, c + 1)[:, :, :n], oracle.fft_inverse(f[c])[:, :, :n]) < 1e-13
    h.close()


@pytest.mark.parametrize("dealias,types", [(0, (100, 200)), (0, (300, 0)), (1, (311, 105))])
def test_rhs_scalars_reactions(oracle, dealias, types):
    """Reacting scalars: add_reaction + dealias_rhs on the 2/3 rule, the
    type-311 reaction with the phase shift (types 1xx-3xx other than 311 get
    no reaction there, as in the reference)."""
    n = 32
    note(f"rhs_scalars reactions dealias={dealias} types={types}")
    h = hit3d_amd.Hit3D(n, n_scalars=2, dealias=dealias)
    nu = 0.02
    pe = (nu / 0.7, nu / 1.3)
    reac = (1.7, 0.6)
    h.set_physics(nu, pe=pe, scalar_type=types)
    h.set_reaction(reac)
    f = solenoidal_fields(oracle, n, 41, dealias, count=5)
    f[3:] *= 0.02   # scalar of order one in physical space
    for c in range(5):
        h.put(FIELDS, c + 1, f[c])
    h.rhs_scalars(True)
    ref = oracle.rhs_scalars(f, 2, pe, types, dealias, n, reac=reac)
    for s in range(2):
        assert rel(h.get(WRK, 4 + s), ref[s]) < 1e-12
    h.close()


def test_force_velocity(oracle):
    n = 32
    note("force_velocity")
    h = hit3d_amd.Hit3D(n)
    h.set_physics(0.01, flow_type=1, kfmax=2, famp=0.5)
    f = solenoidal_fields(oracle, n, 50, 0)
    r = np.stack([oracle.fft_forward(p) for p in phys_random(n, 51, 3)])
    for c in range(3):
        h.put(FIELDS, c + 1, f[c])
        h.put(WRK, c + 1, r[c])
    energy = h.force_velocity()
    rr, e_ref = oracle.force_velocity(f, r, 2, 0.5, n)
    assert abs(energy - e_ref) < 1e-14 * e_ref
    for c in range(3):
        assert rel(h.get(WRK, c + 1), rr[c]) < 1e-14
    h.close()


@pytest.mark.parametrize("dealias", [0, 1])
def test_stats(oracle, dealias):
    n = 32
    note(f"stats dealias={dealias}")
    h = hit3d_amd.Hit3D(n, dealias=dealias)
    f = solenoidal_fields(oracle, n, 60, dealias)
    for c in range(3):
        h.put(FIELDS, c + 1, f[c])
    e, ens = h.stats()
    st = oracle.stats(f, n, oracle.kmax_of(n, dealias), 0.01, 100.0)
    assert rel(e, st["e_spec"]) < 1e-13
    assert abs(ens - st["enstrophy"]) < 1e-12 * st["enstrophy"]
    h.close()


@pytest.mark.parametrize("isp_type", [2, 1])
def test_init_velocity(oracle, isp_type):
    n = 32
    note(f"init_velocity isp={isp_type}")
    h = hit3d_amd.Hit3D(n)
    kmax = oracle.kmax_of(n, 0)
    h.init_velocity(RN1, isp_type, 4, 3.0)
    f, g, _ = oracle.init_velocity(n, RN1, kmax, isp_type, 3.0, 4.0)
    for c in range(3):
        assert rel(h.get(WRK, c + 1), oracle.fft_forward(g[c])) < 1e-13
        assert rel(h.get(FIELDS, c + 1), f[c]) < 1e-12
    h.close()


# ------------------------------------------------------------------ time steps

def _run_steps(oracle, h, cfg, st, steps, fused):
    n = cfg.n
    fov = fos = True
    dt = st.dt
    for it in range(steps):
        c, e, fov, fos = h.step(dt, fov, fos, int_scalars=cfg.n_scalars > 0)
        st = oracle.step(cfg, st)
        assert abs(c - st.courant) < 1e-12 * st.courant, it
        if cfg.flow_type == 1:
            assert abs(e - st.energy) < 1e-12 * st.energy, it
        if cfg.variable_dt:
            dt = oracle.my_dt(dt, c, cfg.n_scalars > 0, cfg.sc)
            assert dt == st.dt
    return st


@pytest.mark.parametrize("case", ["decay23", "forced_ps_scalars"])
def test_trajectory(oracle, case):
    """Several driver steps (main.f90:124-176) against the oracle."""
    if case == "decay23":
        cfg = oracle.Config(n=32, nu=0.01, dt=0.01, dealias=0, flow_type=0)
        steps = 10
    else:
        cfg = oracle.Config(n=16, nu=1 / 400, dt=0.001, dealias=1, flow_type=1, kfmax=2,
                            famp=0.5, variable_dt=True, n_scalars=2, sc=(0.7, 1.0),
                            scalar_type=(0, 0))
        steps = 6
    note(f"trajectory {case}")
    n = cfg.n
    h = hit3d_amd.Hit3D(n, n_scalars=cfg.n_scalars, dealias=cfg.dealias)
    h.set_physics(cfg.nu, flow_type=cfg.flow_type, kfmax=cfg.kfmax, famp=cfg.famp, pe=cfg.pe,
                  scalar_type=cfg.scalar_type)
    h.init_velocity(RN1, 2, 4, 3.0)
    h.dealias_all()
    f0 = np.zeros((3 + cfg.n_scalars, n, n, n + 2))
    for c in range(3):
        f0[c] = h.get(FIELDS, c + 1)
    st = oracle.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt)
    st = _run_steps(oracle, h, cfg, st, steps, fused=False)
    for c in range(3 + cfg.n_scalars):
        assert rel(h.get(FIELDS, c + 1), st.fields[c]) < 1e-11, c
    e, ens = h.stats()
    ref = oracle.stats(st.fields, n, cfg.kmax, cfg.nu, 1 / cfg.nu)
    assert rel(e, ref["e_spec"]) < 1e-11
    h.close()
