"""ORACLE -- test infrastructure only.

A CPU restatement, in numpy (FP64) and C (``randu.c``), of the reference's
pseudospectral time step, in the reference's own array layout:

* physical slot  ``a[z, y, x]`` with ``x`` padded to ``n+2``   (Fortran ``wrk(x,y,z)``)
* spectral slot  ``a[ky, kz, i]`` with ``i = 2*kx + {0: re, 1: im}``
  (Fortran ``wrk(i,kz,ky)``, x_fftw.f90:321-327)

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this module, and only as the checker; the product path
(``hit3d_amd``) never calls it.

Each function cites the reference file:line it follows.  The 3-D FFT is
numpy's pocketfft (an implementation independent of both FFTW and the HIP
kernels); its layout mapping is pinned by single-mode known answers and by a
brute-force DFT in the tests.  Functions whose only check is a closed form or
invariant say so in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

# ----------------------------------------------------------------------------
# C part (RNG, Box-Muller, phase factors) -- randu.c
# ----------------------------------------------------------------------------


class _RanduState(ctypes.Structure):
    _fields_ = [("idum2", ctypes.c_int64), ("iy", ctypes.c_int64),
                ("iv", ctypes.c_int64 * 32)]


_clib = None


def _lib():
    global _clib
    if _clib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "liboracle.so"], cwd=_HERE)
        lib = ctypes.CDLL(path)
        P = ctypes.POINTER
        lib.randu_reset.argtypes = [P(_RanduState)]
        lib.randu_random.argtypes = [P(_RanduState), ctypes.c_double]
        lib.randu_random.restype = ctypes.c_double
        lib.randu_draw.argtypes = [P(_RanduState), ctypes.c_double, ctypes.c_int64,
                                   P(ctypes.c_double)]
        lib.randu_fill.argtypes = lib.randu_draw.argtypes
        lib.randu_fill.restype = ctypes.c_double
        lib.box_muller.argtypes = [ctypes.c_int64, P(ctypes.c_double), P(ctypes.c_double),
                                   P(ctypes.c_double)]
        lib.phase_factors.argtypes = [ctypes.c_int64, P(ctypes.c_double), ctypes.c_double,
                                      P(ctypes.c_double), P(ctypes.c_double)]
        _clib = lib
    return _clib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class RanduGenerator:
    """RANDu.f:12-52 with its SAVEd state (iv, iy, idum2) made explicit."""

    def __init__(self):
        self._s = _RanduState()
        _lib().randu_reset(ctypes.byref(self._s))

    def random(self, idum3: float) -> float:
        return _lib().randu_random(ctypes.byref(self._s), float(idum3))

    def draw(self, seed: float, count: int) -> np.ndarray:
        out = np.empty(count, np.float64)
        _lib().randu_draw(ctypes.byref(self._s), float(seed), count, _dp(out))
        return out

    @property
    def iy(self) -> int:
        return int(self._s.iy)


def box_muller(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """init_velocity.f90:113 with glibc log/sin (elementwise, same shapes)."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    g = np.empty_like(a)
    _lib().box_muller(a.size, _dp(a), _dp(b), _dp(g))
    return g


# ----------------------------------------------------------------------------
# layout helpers
# ----------------------------------------------------------------------------

def to_complex(spec: np.ndarray) -> np.ndarray:
    """Interleaved spectral slot(s) [..., ky, kz, 2*kx+re/im] -> complex [..., ky, kz, kx]."""
    s = np.ascontiguousarray(spec, np.float64)
    return s[..., 0::2] + 1j * s[..., 1::2]


def to_interleaved(c: np.ndarray) -> np.ndarray:
    out = np.empty(c.shape[:-1] + (2 * c.shape[-1],), np.float64)
    out[..., 0::2] = c.real
    out[..., 1::2] = c.imag
    return out


# ----------------------------------------------------------------------------
# x_fftw: wavenumbers, masks, transforms, derivative
# ----------------------------------------------------------------------------

def wavenumbers(n: int):
    """x_fftw.f90:268-296 on one rank: akx(1:n+2), aky(1:n), akz(1:n)."""
    akx = np.empty(n + 2)
    for ix in range(1, n + 2, 2):
        akx[ix - 1] = float((ix - 1) // 2)
        akx[ix] = akx[ix - 1]
    aky = np.empty(n)
    for iy in range(1, n + 1):
        v = float(iy - 1)
        if v > 0.5 * n:
            v -= n
        aky[iy - 1] = v
    akz = aky.copy()
    return akx, aky, akz


def ialias(n: int) -> np.ndarray:
    """x_fftw.f90:298-310: number of components of k above real(n/3).
    Returned as int8 [ky, kz, i]."""
    akx, aky, akz = wavenumbers(n)
    rnx3 = float(n // 3)
    ia = ((np.abs(aky) > rnx3)[:, None, None].astype(np.int8)
          + (np.abs(akz) > rnx3)[None, :, None].astype(np.int8)
          + (np.abs(akx) > rnx3)[None, None, :].astype(np.int8))
    return ia.astype(np.int8)


def kmax_of(n: int, dealias: int) -> int:
    """m_parameters.f90:122-131."""
    if dealias == 0:
        return n // 3
    if dealias == 1:
        return int(np.floor(float(n) / 3.0 * np.sqrt(2.0)))
    raise ValueError("dealias must be 0 or 1")


def _kgrid(n: int):
    """Complex-layout wavenumber broadcasts (kx from akx(1::2))."""
    akx, aky, akz = wavenumbers(n)
    kx = akx[0::2][None, None, :]
    ky = aky[:, None, None]
    kz = akz[None, :, None]
    return kx, ky, kz


def fft_forward(phys: np.ndarray) -> np.ndarray:
    """xFFT3d(+1,n), x_fftw.f90:865-973: unnormalised forward (sign -1) real
    3-D DFT of a[z, y, 0:n]; result in (kx re/im, kz, ky) order."""
    n = phys.shape[0]
    f = np.fft.rfftn(np.ascontiguousarray(phys[:, :, :n]), axes=(0, 1, 2))  # [kz, ky, kx]
    return to_interleaved(np.ascontiguousarray(f.transpose(1, 0, 2)))


def fft_inverse(spec: np.ndarray) -> np.ndarray:
    """xFFT3d(-1,n), x_fftw.f90:976-1082: inverse (sign +1) times norm=1/n^3
    (x_fftw.f90:262,1070).  The imaginary parts of the kx=0 and kx=n/2 terms
    are ignored by the final real transform, as FFTW's c2r does.  The two pad
    entries x=n+1,n+2 are left 0 here (the reference leaves them undefined)."""
    n = spec.shape[0]
    c = to_complex(spec).transpose(1, 0, 2)  # [kz, ky, kx]
    out = np.zeros((n, n, n + 2))
    out[:, :, :n] = np.fft.irfftn(c, s=(n, n, n), axes=(0, 1, 2))
    return out


def x_derivative(spec: np.ndarray, axis: str) -> np.ndarray:
    """x_fftw.f90:637-682: i * k_axis * w."""
    n = spec.shape[0]
    kx, ky, kz = _kgrid(n)
    k = {"x": kx, "y": ky, "z": kz}[axis]
    c = to_complex(spec) * k
    return to_interleaved(1j * c)


# ----------------------------------------------------------------------------
# shells
# ----------------------------------------------------------------------------

def shell_index_f32(k2: np.ndarray) -> np.ndarray:
    """nint(sqrt(real(k2,4))) -- m_stats.f90:218, m_force.f90:40,139,
    init_velocity.f90:171,228 (float32 square root, round half away from 0)."""
    r = np.sqrt(np.asarray(k2, np.float64).astype(np.float32))
    return np.floor(r.astype(np.float64) + 0.5).astype(np.int64)


def shell_index_f64(k2: np.ndarray) -> np.ndarray:
    """nint(sqrt(k2)) in FP64 -- init_velocity.f90:127."""
    return np.floor(np.sqrt(np.asarray(k2, np.float64)) + 0.5).astype(np.int64)


def _k2_interleaved(n: int) -> np.ndarray:
    akx, aky, akz = wavenumbers(n)
    return (akx[None, None, :] ** 2 + aky[:, None, None] ** 2) + akz[None, :, None] ** 2


def shell_spectrum(fields: np.ndarray, n: int, kshells: int) -> np.ndarray:
    """Shell sums of |u|^2 / n^6 over entries i=1..n (Nyquist column
    excluded), kx=0 entries halved; loop order k, j, i.  m_stats.f90:207-228
    (get_e_spec), m_force.f90:133-147, init_velocity.f90:159-180.
    Returns e[0..kshells] (index 0 unused)."""
    akx, _, _ = wavenumbers(n)
    fac = 1.0 / float(n * n * n) ** 2
    k2 = _k2_interleaved(n)[:, :, :n]
    sh = shell_index_f32(k2)
    u = fields[0][:, :, :n]
    v = fields[1][:, :, :n]
    w = fields[2][:, :, :n]
    fac2 = fac * ((u * u + v * v) + w * w)
    fac2 = np.where((akx[:n] == 0.0)[None, None, :], fac2 * 0.5, fac2)
    sel = (sh > 0) & (sh <= kshells)
    e = np.bincount(sh[sel].ravel(), weights=fac2[sel].ravel(), minlength=kshells + 1)
    return e[:kshells + 1]


def shell_spectrum_chunked(fields, n: int, kshells: int, planes: int = 32) -> np.ndarray:
    """shell_spectrum for grids too large for whole-array temporaries: the
    same terms, summed ``planes`` ky-planes at a time (the summation order
    differs, so results agree to rounding).  ``fields`` is a sequence of
    three spectral slots [ky, kz, i]."""
    akx, aky, akz = wavenumbers(n)
    fac = 1.0 / float(n * n * n) ** 2
    half = (akx[:n] == 0.0)[None, None, :]
    e = np.zeros(kshells + 1)
    for k0 in range(0, n, planes):
        k1 = min(n, k0 + planes)
        k2 = (akx[None, None, :n] ** 2 + aky[k0:k1, None, None] ** 2) + akz[None, :, None] ** 2
        sh = shell_index_f32(k2)
        u = fields[0][k0:k1, :, :n]
        v = fields[1][k0:k1, :, :n]
        w = fields[2][k0:k1, :, :n]
        f2 = fac * ((u * u + v * v) + w * w)
        f2 = np.where(half, f2 * 0.5, f2)
        sel = (sh > 0) & (sh <= kshells)
        e += np.bincount(sh[sel].ravel(), weights=f2[sel].ravel(), minlength=kshells + 1)[
            :kshells + 1]
    return e


# ----------------------------------------------------------------------------
# right-hand sides
# ----------------------------------------------------------------------------

def phase_factors(n: int):
    """cos/sin(half*(akx+aky+akz)*dx), rhs_velocity.f90:248-249, glibc."""
    akx, aky, akz = wavenumbers(n)
    ks = np.ascontiguousarray(((akx[0::2][None, None, :] + aky[:, None, None])
                               + akz[None, :, None]).astype(np.float64))
    c = np.empty_like(ks)
    s = np.empty_like(ks)
    dx = 2.0 * np.pi / n
    _lib().phase_factors(ks.size, _dp(ks), dx, _dp(c), _dp(s))
    return c + 1j * s


def courant_number(phys_uvw, n: int, dt: float) -> float:
    """rhs_velocity.f90:31-38."""
    dx = 2.0 * np.pi / n
    a = (np.abs(phys_uvw[0]) + np.abs(phys_uvw[1])) + np.abs(phys_uvw[2])
    return float(np.max(a[:, :, :n])) * dt / dx


def _products(u, v, w):
    return [u * u, u * v, u * w, v * v, v * w, w * w]


def rhs_velocity(fields: np.ndarray, nu: float, dealias: int, n: int):
    """rhs_velocity.f90:1-410 (without the LES hook).  ``fields`` holds at
    least the three spectral velocity slots.  Returns (rhs[3] interleaved,
    physical velocity[3])."""
    kx, ky, kz = _kgrid(n)
    ia = ialias(n)[:, :, 0::2]
    U = [to_complex(fields[c]) for c in range(3)]
    phys = [fft_inverse(fields[c]) for c in range(3)]
    u, v, w = (p[:, :, :n] for p in phys)
    P = [to_complex(fft_forward(_pad(p, n))) for p in _products(u, v, w)]
    lap = -nu * ((kx ** 2 + ky ** 2) + kz ** 2)
    t = [-((kx * P[0] + ky * P[1]) + kz * P[2]),
         -((kx * P[1] + ky * P[3]) + kz * P[4]),
         -((kx * P[2] + ky * P[4]) + kz * P[5])]
    if dealias == 0:
        rhs = [np.where(ia > 0, 0.0, 1j * t[c] + lap * U[c]) for c in range(3)]
        return np.stack([to_interleaved(r) for r in rhs]), np.stack(phys)
    # phase shift + truncation, rhs_velocity.f90:150-410
    e = phase_factors(n)
    ce = np.conj(e)
    rhs = [np.where(ia > 1, 0.0, 0.5j * t[c] + lap * U[c]) for c in range(3)]
    us, vs, ws = (fft_inverse(to_interleaved(U[c] * e))[:, :, :n] for c in range(3))
    Ps = [to_complex(fft_forward(_pad(p, n))) * ce for p in _products(us, vs, ws)]
    rhs[0] = rhs[0] - 0.5j * ((kx * Ps[0] + ky * Ps[1]) + kz * Ps[2])
    rhs[1] = rhs[1] - 0.5j * ((kx * Ps[1] + ky * Ps[3]) + kz * Ps[4])
    rhs[2] = rhs[2] - 0.5j * ((kx * Ps[2] + ky * Ps[4]) + kz * Ps[5])
    rhs = [np.where(ia >= 2, 0.0, r) for r in rhs]
    return np.stack([to_interleaved(r) for r in rhs]), np.stack(phys)


def _pad(p, n):
    out = np.zeros((n, n, n + 2))
    out[:, :, :n] = p
    return out


def rhs_velocity_deviatoric(fields: np.ndarray, nu: float, n: int):
    """The divergence form of rhs_velocity.f90:56-102 (2/3 rule) applied to
    the trace-free product tensor D_ij = u_i u_j - delta_ij |u|^2/3, an
    option of this build: five independent products instead of six, and
    r-hat_i = -i k_j D_ij^ - nu k^2 u-hat_i on ialias == 0 (the mask of
    :104-144), with D_33 = -(D_11 + D_22).  It differs from the divergence
    form by the gradient i k (|u|^2/3)^, which the projection removes.
    Returns (rhs[3] interleaved, physical velocity[3])."""
    kx, ky, kz = _kgrid(n)
    ia = ialias(n)[:, :, 0::2]
    U = [to_complex(fields[c]) for c in range(3)]
    phys = [fft_inverse(fields[c]) for c in range(3)]
    u, v, w = (p[:, :, :n] for p in phys)
    t3 = ((u * u + v * v) + w * w) / 3.0
    D11, D12, D13, D22, D23 = (to_complex(fft_forward(_pad(d, n)))
                               for d in (u * u - t3, u * v, u * w, v * v - t3, v * w))
    D33 = -(D11 + D22)
    lap = -nu * ((kx ** 2 + ky ** 2) + kz ** 2)
    t = [-((kx * D11 + ky * D12) + kz * D13),
         -((kx * D12 + ky * D22) + kz * D23),
         -((kx * D13 + ky * D23) + kz * D33)]
    rhs = [np.where(ia > 0, 0.0, 1j * t[c] + lap * U[c]) for c in range(3)]
    return np.stack([to_interleaved(r) for r in rhs]), np.stack(phys)


def rhs_velocity_rotational(fields: np.ndarray, nu: float, n: int):
    """The rotational form of the velocity RHS with the 2/3 rule, an option
    of this build replacing the divergence form of rhs_velocity.f90:56-102:
    omega = i k x u-hat, the products u x omega in physical space, then
    r-hat = (u x omega)^ - nu k^2 u-hat on ialias == 0, zero elsewhere
    (the mask of :104-144).  It differs from the divergence form by the
    gradient i k (|u|^2/2)^, which the projection removes.  Returns (rhs[3]
    interleaved, physical velocity[3])."""
    kx, ky, kz = _kgrid(n)
    ia = ialias(n)[:, :, 0::2]
    U = [to_complex(fields[c]) for c in range(3)]
    om = [1j * (ky * U[2] - kz * U[1]),
          1j * (kz * U[0] - kx * U[2]),
          1j * (kx * U[1] - ky * U[0])]
    phys = [fft_inverse(fields[c]) for c in range(3)]
    u, v, w = (p[:, :, :n] for p in phys)
    o1, o2, o3 = (fft_inverse(to_interleaved(o))[:, :, :n] for o in om)
    cr = [v * o3 - w * o2, w * o1 - u * o3, u * o2 - v * o1]
    C = [to_complex(fft_forward(_pad(c, n))) for c in cr]
    lap = -nu * ((kx ** 2 + ky ** 2) + kz ** 2)
    rhs = [np.where(ia > 0, 0.0, C[c] + lap * U[c]) for c in range(3)]
    return np.stack([to_interleaved(r) for r in rhs]), np.stack(phys)


def reaction_rate(phi: np.ndarray, rtype: int, rrate: float, mean: float) -> np.ndarray:
    """add_reaction (rhs_scalars.f90:405-456) in physical space: rtype 1 KPP
    r(1-phi^2), 2 symmetric bistable r(1-phi^2)phi, 3 self-adjusting
    bistable r(1-phi^2)(phi-mean)."""
    if rtype == 1:
        return rrate * (1.0 - phi ** 2)
    if rtype == 2:
        return rrate * (1.0 - phi ** 2) * phi
    if rtype == 3:
        return rrate * (1.0 - phi ** 2) * (phi - mean)
    raise ValueError("Unknown reaction rate")


def _cube_mask(n: int, kcut: int) -> np.ndarray:
    kx, ky, kz = _kgrid(n)
    return (np.abs(kx) > kcut) | (np.abs(ky) > kcut) | (np.abs(kz) > kcut)


def rhs_scalars(fields: np.ndarray, n_scalars: int, pe, scalar_type, dealias: int, n: int,
                reac=None):
    """rhs_scalars.f90:48-134 (2/3 rule, with add_reaction :405-456 and
    dealias_rhs :462-496 for types >= 100), :144-271 (phase shift, with the
    type-311 reaction of :280-380) and the mean-gradient source :387-389
    (scalar type 0 -> RHS -= u-hat).  Returns spectral RHS[n_scalars]."""
    reac = [0.0] * n_scalars if reac is None else list(reac)
    nxyz = float(n) ** 3
    kx, ky, kz = _kgrid(n)
    ia = ialias(n)[:, :, 0::2]
    k2 = (kx ** 2 + ky ** 2) + kz ** 2
    U = [to_complex(fields[c]) for c in range(3)]
    out = []
    if dealias == 0:
        u, v, w = (fft_inverse(fields[c])[:, :, :n] for c in range(3))
        for s in range(n_scalars):
            phi = fft_inverse(fields[3 + s])[:, :, :n]
            Q = [to_complex(fft_forward(_pad(phi * q, n))) for q in (u, v, w)]
            conv = (kx * Q[0] + ky * Q[1]) + kz * Q[2]
            r = -1j * conv - pe[s] * k2 * to_complex(fields[3 + s])
            r = np.where(ia > 0, 0.0, r)
            if scalar_type[s] >= 100:
                rate = reaction_rate(phi, scalar_type[s] // 100, reac[s],
                                     fields[3 + s][0, 0, 0] / nxyz)
                r = r + to_complex(fft_forward(_pad(rate, n)))
                r = np.where(_cube_mask(n, kmax_of(n, 0)), 0.0, r)
            out.append(r)
    else:
        e = phase_factors(n)
        ce = np.conj(e)
        us, vs, ws = (fft_inverse(to_interleaved(U[c] * e))[:, :, :n] for c in range(3))
        u, v, w = (fft_inverse(fields[c])[:, :, :n] for c in range(3))
        for s in range(n_scalars):
            F = to_complex(fields[3 + s])
            phs = fft_inverse(to_interleaved(F * e))[:, :, :n]
            Qs = [to_complex(fft_forward(_pad(phs * q, n))) * ce for q in (us, vs, ws)]
            half1 = -0.5j * ((kx * Qs[0] + ky * Qs[1]) + kz * Qs[2])
            phi = fft_inverse(fields[3 + s])[:, :, :n]
            Q = [to_complex(fft_forward(_pad(phi * q, n))) for q in (u, v, w)]
            half2 = -0.5j * ((kx * Q[0] + ky * Q[1]) + kz * Q[2]) - pe[s] * k2 * F
            r = np.where(ia > 1, 0.0, half1 + half2)
            if scalar_type[s] > 300:
                if scalar_type[s] != 311:
                    raise ValueError("reaction type not supported with phase shifting")
                trunc = np.where(_cube_mask(n, n // 4), 0.0, F)
                ph = fft_inverse(to_interleaved(trunc))[:, :, :n]
                rate = reaction_rate(ph, 3, reac[s], fields[3 + s][0, 0, 0] / nxyz)
                r = r + to_complex(fft_forward(_pad(rate, n)))
            out.append(r)
    for s in range(n_scalars):
        if scalar_type[s] == 0:
            out[s] = out[s] - U[0]
    if not out:
        return np.zeros((0, n, n, n + 2))
    return np.stack([to_interleaved(r) for r in out])


# ----------------------------------------------------------------------------
# forcing, update, projection, masks, dt
# ----------------------------------------------------------------------------

def _strain(fields, n):
    """S_ij (complex) of m_les.f90:1154-1196: S11, S12, S13, S22, S23, S33."""
    kx, ky, kz = _kgrid(n)
    U = [to_complex(fields[c]) for c in range(3)]
    return [1j * kx * U[0],
            0.5j * (ky * U[0] + kx * U[1]),
            0.5j * (kz * U[0] + kx * U[2]),
            1j * ky * U[1],
            0.5j * (kz * U[1] + ky * U[2]),
            1j * kz * U[2]]


def turb_visc_smag(fields: np.ndarray, n: int, c_smag: float = 0.18) -> np.ndarray:
    """les_get_turb_visc_smag (m_les.f90:1142-1215): (c_s dx)^2 sqrt(2 S_ij S_ij),
    in physical space (n, n, n), les_delta = dx (m_les.f90:86)."""
    S = [fft_inverse(to_interleaved(s))[:, :, :n] for s in _strain(fields, n)]
    tv = S[0] ** 2
    tv = tv + 2.0 * S[1] ** 2
    tv = tv + 2.0 * S[2] ** 2
    tv = tv + S[3] ** 2
    tv = tv + 2.0 * S[4] ** 2
    tv = tv + S[5] ** 2
    dx = 2.0 * np.pi / n
    return np.sqrt(2.0 * tv) * (c_smag * dx) ** 2


def les_rhs_velocity(fields: np.ndarray, rhs: np.ndarray, tv: np.ndarray, n: int) -> np.ndarray:
    """les_rhsv_turb_visc (m_les.f90:902-1018) then the ialias != 0 mask of
    les_rhs_velocity (:682-690): rhs_i += -i k_j tau_ij, tau_ij = -2 nu_t S_ij."""
    kx, ky, kz = _kgrid(n)
    ia = ialias(n)[:, :, 0::2]
    S = _strain(fields, n)
    tau = []
    for s in S:
        ph = fft_inverse(to_interleaved(s))[:, :, :n]
        tau.append(to_complex(fft_forward(_pad(-2.0 * tv * ph, n))))
    r = [to_complex(rhs[c]) for c in range(3)]
    keep = ia == 0
    r[0] = np.where(keep, r[0] - 1j * ((kx * tau[0] + ky * tau[1]) + kz * tau[2]), r[0])
    r[1] = np.where(keep, r[1] - 1j * ((kx * tau[1] + ky * tau[3]) + kz * tau[4]), r[1])
    r[2] = np.where(keep, r[2] - 1j * ((kx * tau[2] + ky * tau[4]) + kz * tau[5]), r[2])
    r = [np.where(keep, x, 0.0) for x in r]
    return np.stack([to_interleaved(x) for x in r])


def velocity_gradients(fields: np.ndarray, n: int):
    """du_i/dx_j in physical space, i-major (11, 12, 13, 21, ..., 33):
    x_derivative then the inverse transform."""
    return [fft_inverse(x_derivative(fields[c], ax))[:, :, :n] for c in range(3) for ax in ("x", "y", "z")]


def stress_divergence(rhs: np.ndarray, tau6, n: int) -> np.ndarray:
    """rhs_i += -i k_j tau_ij (tau physical: 11, 12, 13, 22, 23, 33) on
    ialias == 0 entries, others zeroed -- the stress divergence and mask of
    les_rhsv_turb_visc / les_rhs_velocity (m_les.f90:937-1015, :682-690)
    for an arbitrary stress."""
    kx, ky, kz = _kgrid(n)
    ia = ialias(n)[:, :, 0::2]
    tau = [to_complex(fft_forward(_pad(t, n))) for t in tau6]
    r = [to_complex(rhs[c]) for c in range(3)]
    keep = ia == 0
    r[0] = np.where(keep, r[0] - 1j * ((kx * tau[0] + ky * tau[1]) + kz * tau[2]), r[0])
    r[1] = np.where(keep, r[1] - 1j * ((kx * tau[1] + ky * tau[3]) + kz * tau[4]), r[1])
    r[2] = np.where(keep, r[2] - 1j * ((kx * tau[2] + ky * tau[4]) + kz * tau[5]), r[2])
    r = [np.where(keep, x, 0.0) for x in r]
    return np.stack([to_interleaved(x) for x in r])


def les_rhs_scalar(field_s: np.ndarray, rhs_s: np.ndarray, tv: np.ndarray, n: int) -> np.ndarray:
    """m_les_rhss_turb_visc1 (m_les.f90:1041-1091): rhs += sum_dir
    d_dir(nu_t d_dir phi) on entries with ialias == 0."""
    ia = ialias(n)
    acc = np.zeros_like(field_s)
    for ax in ("x", "y", "z"):
        d = fft_inverse(x_derivative(field_s, ax))[:, :, :n]
        acc = acc + x_derivative(fft_forward(_pad(d * tv, n)), ax)
    return np.where(ia == 0, rhs_s + acc, rhs_s)


def forced_mask(n: int, kfmax: int) -> np.ndarray:
    """m_force.f90:37-78: entries i=1..n with 0 < shell <= kfmax (interleaved)."""
    k2 = _k2_interleaved(n)
    sh = shell_index_f32(k2)
    m = (sh > 0) & (sh <= kfmax)
    m[:, :, n:] = False
    return m


def force_velocity(fields: np.ndarray, rhs: np.ndarray, kfmax: int, famp: float, n: int):
    """m_force.f90:101-182.  Returns (rhs with forcing added, band energy)."""
    e = shell_spectrum(fields, n, kfmax)
    energy = float(np.sum(e[1:kfmax + 1]))
    fac = famp / energy
    m = forced_mask(n, kfmax)
    out = rhs.copy()
    for c in range(3):
        out[c] = np.where(m, out[c] + fac * fields[c], out[c])
    return out, energy


def advance(fields, rhs, rhs_old, dt: float, first_order: bool):
    """main.f90:137-146 (scalars) and :156-164 (velocity): Euler on the first
    step, else AB2.  Returns (fields, rhs_old)."""
    if first_order:
        return fields + dt * rhs, rhs.copy()
    return fields + dt * (1.5 * rhs - 0.5 * rhs_old), rhs.copy()


def pressure(fields: np.ndarray, n: int) -> np.ndarray:
    """pressure.f90:13-47: u -= k (k.u)/k^2, k^2=0 -> 9e20."""
    kx, ky, kz = _kgrid(n)
    U = [to_complex(fields[c]) for c in range(3)]
    div = (kx * U[0] + ky * U[1]) + kz * U[2]
    lap = (kx ** 2 + ky ** 2) + kz ** 2
    lap = np.where(lap == 0.0, 9e20, lap)
    p = div.real / lap + 1j * (div.imag / lap)
    out = fields.copy()
    out[0] = to_interleaved(U[0] - p * kx)
    out[1] = to_interleaved(U[1] - p * ky)
    out[2] = to_interleaved(U[2] - p * kz)
    return out


def dealias_all(fields: np.ndarray, kmax: int, n: int) -> np.ndarray:
    """dealias_all.f90:19-32: cube truncation at kmax."""
    akx, aky, akz = wavenumbers(n)
    a = float(kmax)
    m = ((np.abs(akx) > a)[None, None, :] | (np.abs(aky) > a)[:, None, None]
         | (np.abs(akz) > a)[None, :, None])
    return np.where(m[None], 0.0, fields)


def my_dt(dt: float, courant: float, int_scalars: bool = False, sc=()) -> float:
    """my_dt.f90:8-30 (flow_type <= 1)."""
    courant_max = 0.1
    if courant < courant_max:
        dt = dt * 1.01
        dt = min(dt, 0.01)
    else:
        dt = dt / 1.05
    if int_scalars:
        for s in sc:
            dt = min(dt, dt * courant_max / courant / s)
    return dt


# ----------------------------------------------------------------------------
# statistics
# ----------------------------------------------------------------------------

def stats(fields: np.ndarray, n: int, kmax: int, nu: float, re: float):
    """m_stats.f90:99-148 (stat_velocity) with get_e_spec (:195-235) and
    get_gradient_statistics (:241-289).  Returns a dict."""
    e = shell_spectrum(fields, n, kmax)
    e_spec = e[1:kmax + 1]
    energy = float(np.sum(e_spec))
    ks = np.arange(1, kmax + 1, dtype=np.float64)
    diss = e_spec * ks ** 2 * 2.0 * nu
    eps_v = float(np.sum(diss))
    # vorticity, m_stats.f90:260-285
    d = {}
    for (c, ax) in ((2, "y"), (2, "x"), (1, "z"), (1, "x"), (0, "z"), (0, "y")):
        d[(c, ax)] = x_derivative(fields[c], ax)
    om3 = d[(1, "x")] - d[(0, "y")]
    om2 = d[(0, "z")] - d[(2, "x")]
    om1 = d[(2, "y")] - d[(1, "z")]
    w1, w2, w3 = (fft_inverse(o)[:, :, :n] for o in (om1, om2, om3))
    enstrophy = float(np.sum((w1 ** 2 + w2 ** 2) + w3 ** 2)) / float(n ** 3)
    eta = (nu ** 3 / eps_v) ** 0.25
    uvar = 2.0 / 3.0 * energy
    x_length = np.pi / 2.0 * float(np.sum(e_spec / ks)) / uvar
    lam = np.sqrt(15.0 * uvar * nu / eps_v)
    return dict(e_spec=e_spec, diss_spec=diss, energy=energy, eps_v=eps_v, eta=eta,
                etakmax=eta * kmax, enstrophy=enstrophy,
                re_lambda=uvar * np.sqrt(15.0 / eps_v * re),
                x_length=x_length, lambda_=lam,
                re_lambda1=np.sqrt(uvar) * lam / nu, tau_e=x_length / np.sqrt(uvar))


def scalar_stats(phi: np.ndarray, n: int, kmax: int, pe: float) -> dict:
    """stat_scalars (m_stats.f90:295-366) with get_scalar_spectra (:371-437)
    for one spectral scalar field phi.  The variance spectrum is the FP64
    shell sum of |phi|^2/n^6 over i=1..n with kx=0 halved (same loop as
    get_e_spec), doubled (:418); dissipation 2 pe <|grad phi|^2> from the
    three derivatives brought to physical space (:320-338); mean, variance,
    min and max from phi in physical space (:341-360)."""
    z = np.zeros_like(phi)
    es = 2.0 * shell_spectrum(np.stack([phi, z, z]), n, kmax)[1:kmax + 1]
    fac = 1.0 / float(n * n * n)
    g = [fft_inverse(x_derivative(phi, ax))[:, :, :n] for ax in ("x", "y", "z")]
    w4 = (g[0] ** 2 + g[1] ** 2) + g[2] ** 2
    sc_diss = 2.0 * pe * float(np.sum(w4)) * fac
    p = fft_inverse(phi)[:, :, :n]
    mean = float(np.sum(p)) * fac
    var = float(np.sum(p ** 2)) * fac - mean ** 2
    return dict(sc_diss=sc_diss, mean=mean, variance=var, sc_min=float(p.min()),
                sc_max=float(p.max()), es_spec=es)


def skewness(fields: np.ndarray, n: int) -> float:
    """Mean over i of <(du_i/dx_i)^3>/<(du_i/dx_i)^2>^(3/2) (a build statistic)."""
    out = 0.0
    for c, ax in ((0, "x"), (1, "y"), (2, "z")):
        d = fft_inverse(x_derivative(fields[c], ax))[:, :, :n]
        out += float(np.mean(d ** 3)) / float(np.mean(d ** 2)) ** 1.5
    return out / 3.0


def divergence(fields: np.ndarray, n: int):
    """pressure.f90:53-93: min and max of du/dx + dv/dy + dw/dz, each
    derivative taken spectrally and brought to physical space separately."""
    d = [fft_inverse(x_derivative(fields[c], ax))[:, :, :n] for c, ax in ((0, "x"), (1, "y"), (2, "z"))]
    s = (d[0] + d[1]) + d[2]
    return float(s.min()), float(s.max())


# ----------------------------------------------------------------------------
# Lagrangian particles
# ----------------------------------------------------------------------------

def particles_lattice(nptot: int, n: int):
    """particles_init (m_particles.f90:213-262,300-327) without a <run>.pt
    file, one rank: nptot rounded to the nearest cube, a cell-centred
    lattice over [0, 2 pi)^3, converted to grid units x/(2 pi/n) + 1.
    Returns (nptot, xyzp[np, 3])."""
    npthird = int(float(nptot) ** (1.0 / 3.0))
    if nptot - npthird ** 3 >= (npthird + 1) ** 3 - nptot:
        npthird += 1
    nptot = min(npthird ** 3, 2 ** 30)
    sctmp = 2.0 * np.pi / float(npthird)
    i = np.arange(npthird, dtype=np.float64)
    kk, jj, ii = np.meshgrid(i, i, i, indexing="ij")
    xyz = np.stack([(ii.ravel() + 0.5) * sctmp, (jj.ravel() + 0.5) * sctmp,
                    (kk.ravel() + 0.5) * sctmp], axis=1)
    two_pi = 2.0 * np.pi
    bad = (xyz >= two_pi) | (xyz < 0.0)
    xyz = np.where(bad, xyz - two_pi * np.floor(xyz / two_pi), xyz)
    return nptot, xyz / (two_pi / float(n)) + 1.0


def _cint_r(x):
    """cint's weight vector R(x), m_particles.f90:1496-1502."""
    s = (x - 1.0) * (x - 2.0)
    r1 = -x * s
    r2 = 3.0 * (x + 1.0) * s
    s = (x + 1.0) * x
    r3 = -3.0 * s * (x - 2.0)
    r4 = s * (x - 1.0)
    return r1, r2, r3, r4


def _cint(r, p1, p2, p3, p4):
    """m_particles.f90:1505: (r1 p1 + r2 p2 + r3 p3 + r4 p4) / 6."""
    return (r[0] * p1 + r[1] * p2 + r[2] * p3 + r[3] * p4) / 6.0


def particles_move(vel, xyzp: np.ndarray, uvwp: np.ndarray, dt: float, n: int, scheme: int):
    """particles_interpolate_trilinear (m_particles.f90:815-1009, scheme 0)
    or particles_interpolate_cint (:1017-1230 with cint_3d :1451-1480,
    scheme 1), then the one-rank slab update (:1236-1300: z >= n+1 -> z-n,
    z < 1 -> z+n).  vel = three physical slots a[z, y, x]; xyzp/uvwp[np, 3]
    in grid units.  Returns new (xyzp, uvwp); particles keep their order."""
    dx = 2.0 * np.pi / float(n)
    x, y, z = xyzp[:, 0], xyzp[:, 1], xyzp[:, 2]
    i = np.floor(x).astype(np.int64)
    j = np.floor(y).astype(np.int64)
    k = np.floor(z).astype(np.int64)
    xp, yp, zp = x - i, y - j, z - k
    f = [np.asarray(v)[:, :, :n] for v in vel]

    def at(c, ii, jj, kk):  # 1-based fields(ii, jj, kk, c), periodic
        return f[c][(kk - 1) % n, (jj - 1) % n, (ii - 1) % n]

    up = []
    if scheme == 0:
        c1 = (1.0 - xp) * (1.0 - yp) * (1.0 - zp)
        c2 = (xp) * (1.0 - yp) * (1.0 - zp)
        c3 = (1.0 - xp) * (yp) * (1.0 - zp)
        c4 = (xp) * (yp) * (1.0 - zp)
        c5 = (1.0 - xp) * (1.0 - yp) * (zp)
        c6 = (xp) * (1.0 - yp) * (zp)
        c7 = (1.0 - xp) * (yp) * (zp)
        c8 = (xp) * (yp) * (zp)
        i1 = np.where(i == n, 1, i + 1)
        j1 = np.where(j == n, 1, j + 1)
        k1 = np.where(k == n, 1, k + 1)   # one rank: the next slab is this one
        for c in range(3):
            s = c1 * at(c, i, j, k) + c2 * at(c, i1, j, k) + c3 * at(c, i, j1, k) + \
                c4 * at(c, i1, j1, k)
            s = s + c5 * at(c, i, j, k1) + c6 * at(c, i1, j, k1) + c7 * at(c, i, j1, k1) + \
                c8 * at(c, i1, j1, k1)
            up.append(s)
    else:
        rx, ry, rz = _cint_r(xp), _cint_r(yp), _cint_r(zp)
        for c in range(3):
            t1 = []
            for kk in range(4):
                t2 = []
                for jj in range(4):
                    pts = [at(c, i - 2 + ii + 1, j - 2 + jj + 1, k - 2 + kk + 1) for ii in range(4)]
                    t2.append(_cint(rx, *pts))
                t1.append(_cint(ry, *t2))
            up.append(_cint(rz, *t1))
    u, v, w = up
    first = uvwp[:, 0] == -999.0
    xe = xp + dt * u / dx
    ye = yp + dt * v / dx
    ze = zp + dt * w / dx
    xa = xp + (1.50 * u - 0.50 * uvwp[:, 0]) * dt / dx
    ya = yp + (1.50 * v - 0.50 * uvwp[:, 1]) * dt / dx
    za = zp + (1.50 * w - 0.50 * uvwp[:, 2]) * dt / dx
    xp = np.where(first, xe, xa)
    yp = np.where(first, ye, ya)
    zp = np.where(first, ze, za)

    def wrap(fr, cell, periodic):
        hi = fr > 1.0
        fr = np.where(hi, fr - 1.0, fr)
        cell = np.where(hi, cell + 1, cell)
        if periodic:
            cell = np.where(hi & (cell > n), cell - n, cell)
        lo = fr < 0.0
        fr = np.where(lo, fr + 1.0, fr)
        cell = np.where(lo, cell - 1, cell)
        if periodic:
            cell = np.where(lo & (cell < 1), cell + n, cell)
        return fr, cell

    xp, i = wrap(xp, i, True)
    yp, j = wrap(yp, j, True)
    zp, k = wrap(zp, k, False)
    zn = zp + k.astype(np.float64)
    zn = np.where(zn >= float(n + 1), zn - float(n), zn)
    zn = np.where(zn < 1.0, zn + float(n), zn)
    new_xyz = np.stack([xp + i.astype(np.float64), yp + j.astype(np.float64), zn], axis=1)
    return new_xyz, np.stack([u, v, w], axis=1)


# ----------------------------------------------------------------------------
# seeded initial velocity
# ----------------------------------------------------------------------------

def target_spectrum(kmax: int, isp_type: int, peak: float, ir_exp: float) -> np.ndarray:
    """init_velocity.f90:193-220 (types 0, 1, 3) plus type 2,
    E ~ k^m exp(-(m/2)(k/k0)^2), which the reference does not have."""
    e = np.empty(kmax)
    for k in range(1, kmax + 1):
        wmag = float(k)
        ratio = wmag / peak
        if isp_type == 0:
            e[k - 1] = wmag ** (-5.0 / 3.0)
        elif isp_type == 1:
            e[k - 1] = ratio ** 3 / peak * np.exp(-3.0 * ratio)
        elif isp_type == 2:
            e[k - 1] = wmag ** ir_exp * np.exp(-0.5 * ir_exp * ratio * ratio)
        elif isp_type == 3:
            fac = 2.0 * np.pi * ratio
            e[k - 1] = fac ** 4 / (1.0 + fac ** 2) ** 3
        else:
            raise ValueError("wrong initial spectrum type")
    return e / np.sum(e)


def init_velocity_uniforms(n: int, rn1: float, gen: RanduGenerator | None = None):
    """init_velocity.f90:67-103 on one rank: random(-RN1), then the six
    slots wrk(1:n+2,1:n,1:n,1..6) filled in memory order."""
    gen = gen or RanduGenerator()
    gen.random(-rn1)
    return gen.draw(rn1, 6 * (n + 2) * n * n).reshape(6, n, n, n + 2), gen


def init_velocity(n: int, rn1: float, kmax: int, isp_type: int, peak: float, ir_exp: float):
    """init_velocity.f90:53-237.  Returns (fields[3] spectral, gaussian[3]
    physical, e_spec_before_rescale)."""
    uni, _ = init_velocity_uniforms(n, rn1)
    g = box_muller(uni[0:3], uni[3:6])
    f, e = init_velocity_from_gaussian(g, n, kmax, isp_type, peak, ir_exp)
    return f, g, e


def init_velocity_from_gaussian(g: np.ndarray, n: int, kmax: int, isp_type: int, peak: float,
                                ir_exp: float):
    """The post-RNG stage of init_velocity (init_velocity.f90:116-237): forward
    transforms of the three physical slots g, the curl i k x W on shells
    1..kmax (FP64 shell index, :123-147), then each shell rescaled to the
    target spectrum (:153-237).  Returns (fields[3] spectral,
    e_spec_before_rescale)."""
    W = [to_complex(fft_forward(g[c])) for c in range(3)]
    kx, ky, kz = _kgrid(n)
    sh = shell_index_f64((kx ** 2 + ky ** 2) + kz ** 2)
    keep = (sh > 0) & (sh <= kmax)
    curl = [1j * (ky * W[2] - kz * W[1]),
            1j * (kz * W[0] - kx * W[2]),
            1j * (kx * W[1] - ky * W[0])]
    f = np.stack([to_interleaved(np.where(keep, c, 0.0)) for c in curl])
    e = shell_spectrum(f, n, kmax)[1:]
    tgt = target_spectrum(kmax, isp_type, peak, ir_exp)
    sh32 = shell_index_f32(_k2_interleaved(n))
    ok = (sh32 > 0) & (sh32 <= kmax)
    e_at = np.where(ok, e[np.clip(sh32, 1, kmax) - 1], 0.0)
    t_at = np.where(ok, tgt[np.clip(sh32, 1, kmax) - 1], 0.0)
    ok &= e_at > 0.0
    scale = np.where(ok, np.sqrt(np.where(ok, t_at / np.where(ok, e_at, 1.0), 0.0)), 0.0)
    f = np.where(ok[None], f * scale[None], 0.0)
    return f, e


def scalar_target_spectrum(kmax: int, sc_type: int, peak: float) -> np.ndarray:
    """init_scalars.f90:170-196: type 0 k^-5/3, 1 and 3 exponential, 2 von
    Karman, normalised to unit sum."""
    e = np.empty(kmax)
    for k in range(1, kmax + 1):
        wmag = float(k)
        ratio = wmag / peak
        if sc_type == 0:
            e[k - 1] = wmag ** (-5.0 / 3.0)
        elif sc_type in (1, 3):
            e[k - 1] = ratio ** 3 / peak * np.exp(-3.0 * ratio)
        elif sc_type == 2:
            fac = 2.0 * np.pi * ratio
            e[k - 1] = fac ** 4 / (1.0 + fac ** 2) ** 3
        else:
            raise ValueError("INIT_SCALARS: WRONG INITIAL SPECTRUM TYPE")
    return e / np.sum(e)


def scalar_space(n: int, ic_type: int, peak: float) -> np.ndarray:
    """init_scalar_space (init_scalars.f90:278-383), physical slot."""
    dx = 2.0 * np.pi / n
    out = np.zeros((n, n, n + 2))
    if ic_type == 11:
        h = max(2.0 * dx, 2.0 * np.pi / peak)
        zloc = np.arange(n) * dx
        v = np.tanh((zloc - np.pi * 0.5) / h) - np.tanh((zloc - np.pi * 1.5) / h) - 1.0
        out[:, :, :n] = v[:, None, None]
    elif ic_type == 12:
        h = max(8.0 * dx, np.pi / 8.0)
        xx = np.arange(n) * dx
        v = (np.tanh((xx - np.pi * 0.25) / h) - np.tanh((xx - np.pi * 0.75) / h)
             + np.tanh((xx - np.pi * 1.25) / h) - np.tanh((xx - np.pi * 1.75) / h))
        out[:, :, :n] = (v - 1.0)[None, None, :]
    elif ic_type == 13:
        pk = min(peak, float(n // 8))
        out[:, :, :n] = np.sin(pk * dx * np.arange(n))[None, None, :]
    else:
        raise ValueError("INIT_SCALARS: UNEXPECTED SCALAR TYPE")
    return out


def init_scalars(n: int, kmax: int, scalar_type, peak_sc, rn2: float, gen: RanduGenerator):
    """init_scalars (init_scalars.f90:6-40) with init_scalar_spectrum
    (:84-270) and init_scalar_space (:278-383) on one rank.  gen carries the
    generator state left by the velocity initialisation.  Returns the
    spectral scalar fields [n_scalars]."""
    out = []
    sh32 = shell_index_f32(_k2_interleaved(n))
    for s, st in enumerate(scalar_type):
        ic = st - (st // 100) * 100
        if ic == 0:
            f = np.zeros((n, n, n + 2))
        elif ic < 10:
            gen.random(-rn2)
            w = gen.draw(rn2, 2 * (n + 2) * n * n).reshape(2, n, n, n + 2)
            g = fft_forward(box_muller(w[0], w[1]))
            e = shell_spectrum(np.stack([g, np.zeros_like(g), np.zeros_like(g)]), n, kmax)[1:]
            tgt = scalar_target_spectrum(kmax, st, peak_sc[s])
            ok = (sh32 > 0) & (sh32 <= kmax)
            e_at = np.where(ok, e[np.clip(sh32, 1, kmax) - 1], 0.0)
            t_at = np.where(ok, tgt[np.clip(sh32, 1, kmax) - 1], 0.0)
            ok &= e_at > 0.0
            scale = np.sqrt(np.where(ok, t_at / np.where(ok, e_at, 1.0), 0.0))
            f = np.where(ok, g * scale, 0.0)
            if st == 3:
                ph = fft_inverse(f)[:, :, :n]
                f = fft_forward(_pad(np.copysign(1.0, ph) * 0.9, n))
                f = np.where((sh32 == 0) | (sh32 >= kmax * 2 // 3), 0.0, f)
        else:
            f = fft_forward(scalar_space(n, ic, peak_sc[s]))
            if ic in (11, 12):
                f[0, 0, 0] = 0.0
        out.append(f)
    ia = ialias(n)
    return [np.where(ia > 0, 0.0, f) for f in out]


def taylor_green(n: int) -> np.ndarray:
    """init_velocity.f90:252-293."""
    dx = 2.0 * np.pi / n
    x = np.arange(n) * dx
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    u = np.sin(xx) * np.cos(yy) * np.cos(zz)
    v = -np.cos(xx) * np.sin(yy) * np.cos(zz)
    return np.stack([fft_forward(_pad(u, n)), fft_forward(_pad(v, n)),
                     fft_forward(np.zeros((n, n, n + 2)))])


# ----------------------------------------------------------------------------
# one time step of the driver loop
# ----------------------------------------------------------------------------

@dataclass
class Config:
    n: int
    nu: float
    dt: float
    dealias: int = 0
    flow_type: int = 0
    kfmax: int = 2
    famp: float = 0.5
    variable_dt: bool = False
    n_scalars: int = 0
    sc: tuple = ()
    scalar_type: tuple = ()
    re: float = 0.0
    les_model: int = 0
    c_smag: float = 0.18
    rotational: bool = False   # u x omega form of the velocity RHS (2/3 rule only)
    deviatoric: bool = False   # trace-free product tensor (2/3 rule only)

    @property
    def kmax(self):
        return kmax_of(self.n, self.dealias)

    @property
    def pe(self):
        return tuple(self.nu / s for s in self.sc)


@dataclass
class State:
    fields: np.ndarray
    rhs_old: np.ndarray
    fov: bool = True
    fos: bool = True
    time: float = 0.0
    dt: float = 0.0
    courant: float = 0.0
    energy: float = 0.0
    extra: dict = field(default_factory=dict)


def step(cfg: Config, st: State) -> State:
    """main.f90:124-176 for one hydro task: rhs_scalars, scalar update,
    rhs_velocity, force_velocity, velocity update, pressure, TIME+=DT, my_dt."""
    n = cfg.n
    fields = st.fields.copy()
    rhs_old = st.rhs_old.copy()
    ns = cfg.n_scalars
    tv = turb_visc_smag(fields, n, cfg.c_smag) if cfg.les_model == 1 else None
    if ns > 0:
        rs = rhs_scalars(fields, ns, cfg.pe, cfg.scalar_type, cfg.dealias, n)
        if tv is not None:
            rs = np.stack([les_rhs_scalar(fields[3 + s], rs[s], tv, n) for s in range(ns)])
        new, old = advance(fields[3:3 + ns], rs, rhs_old[3:3 + ns], st.dt, st.fos)
        fields[3:3 + ns] = new
        rhs_old[3:3 + ns] = old
    if cfg.rotational or cfg.deviatoric:
        if cfg.dealias != 0:
            raise ValueError("the rotational and deviatoric forms are built for the 2/3 rule")
        form = rhs_velocity_rotational if cfg.rotational else rhs_velocity_deviatoric
        rhs, phys = form(fields, cfg.nu, n)
    else:
        rhs, phys = rhs_velocity(fields, cfg.nu, cfg.dealias, n)
    if tv is not None:
        rhs = les_rhs_velocity(fields, rhs, tv, n)
    courant = courant_number(phys, n, st.dt)
    energy = st.energy
    if cfg.flow_type == 1:
        rhs, energy = force_velocity(fields, rhs, cfg.kfmax, cfg.famp, n)
    new, old = advance(fields[0:3], rhs, rhs_old[0:3], st.dt, st.fov)
    fields[0:3] = new
    rhs_old[0:3] = old
    fields = pressure(fields, n)
    time = st.time + st.dt
    dt = st.dt
    if cfg.variable_dt:
        dt = my_dt(dt, courant, ns > 0, cfg.sc)
    return State(fields=fields, rhs_old=rhs_old, fov=False, fos=st.fos if ns == 0 else False,
                 time=time, dt=dt, courant=courant, energy=energy)
