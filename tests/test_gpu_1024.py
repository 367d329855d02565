"""GPU checks at 1024^3, the largest grid of the measured configurations
(one FP64 slot is 8.6 GB; the transforms and the fused step run kernels of
their own at this size: radix-16 x pass, 1024-point columns).

* Transforms: a single Fourier mode goes to one coefficient N^3 amp / 2 and
  nothing else (1e-11 N^3), and back to the same field (1e-13 relative).
* Fused step: two steps (Euler, then AB2) from one initial field, fused and
  unfused, give the same Courant number and band energy (1e-13 relative)
  and the same E(k) and enstrophy from the device statistics (1e-12
  relative); the divergence after the fused steps is at roundoff (the field
  of the unfused run is restored from a device copy, so no slot crosses
  PCIe).
"""
import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS, WRK

pytestmark = pytest.mark.gpu

N = 1024
RN1 = 67748937.0


def note(msg):
    print(msg, flush=True)


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def test_fft_single_mode_roundtrip_1024():
    note("fft single mode n=1024")
    h = hit3d_amd.Hit3D(N, dealias=0)
    try:
        kx, ky, kz, amp = 301, -77, 455, 0.5
        x = 2 * np.pi * np.arange(N) / N
        a = np.zeros((N, N, N + 2))
        # u(x, y, z) = amp cos(kx x + ky y + kz z), array index (z, y, x)
        a[:, :, :N] = amp * np.cos((kz * x)[:, None, None] + (ky * x)[None, :, None]
                                   + (kx * x)[None, None, :])
        h.put(WRK, 1, a)
        h.xFFT3d(1, 1)
        s = h.get(WRK, 1)
        # spectral layout (ky, kz, kx re/im); forward sign -1, unnormalised
        j, k = kz % N, ky % N
        assert abs(s[k, j, 2 * kx] - N ** 3 * amp / 2) < 1e-9 * N ** 3
        assert abs(s[k, j, 2 * kx + 1]) < 1e-9 * N ** 3
        s[k, j, 2 * kx:2 * kx + 2] = 0.0
        assert np.abs(s).max() < 1e-11 * N ** 3
        del s
        h.xFFT3d(-1, 1)
        back = h.get(WRK, 1)
        assert rel(back[:, :, :N], a[:, :, :N]) < 1e-13
    finally:
        h.close()


def test_fused_step_1024_matches_unfused():
    note("fused vs unfused step n=1024")
    nu = 1 / 1600.0
    # three scalar slots that the velocity-only step leaves alone hold the
    # initial velocity on the device
    h = hit3d_amd.Hit3D(N, n_scalars=3, dealias=0)
    try:
        h.set_physics(nu, flow_type=1, kfmax=2, famp=0.5, pe=(nu, nu, nu), scalar_type=(0, 0, 0))
        h.init_velocity(RN1, 2, 4, 3.0)
        h.dealias_all()
        for c in range(3):
            h.copy(FIELDS, 4 + c, FIELDS, 1 + c)
        res = {}
        for fused in (True, False):
            for c in range(3):
                h.copy(FIELDS, 1 + c, FIELDS, 4 + c)
            h.set_fusion(fused)
            fov, steps = True, []
            for _ in range(2):
                cr, en, fov, _ = h.step(2.5e-4, fov, True)
                steps.append((cr, en))
            e, ens = h.stats()
            res[fused] = (steps, e, ens, h.divergence())
            note(f"fused={fused}: courant {steps[-1][0]:.6f} energy {steps[-1][1]:.6e} "
                 f"enstrophy {ens:.6e} divergence {res[fused][3]}")
        (s1, e1, z1, d1), (s2, e2, z2, d2) = res[True], res[False]
        for (c1, n1), (c2, n2) in zip(s1, s2):
            assert abs(c1 - c2) <= 1e-13 * c2 and abs(n1 - n2) <= 1e-13 * n2
        assert rel(e1, e2) < 1e-12
        assert abs(z1 - z2) <= 1e-12 * z2
        assert max(abs(d1[0]), abs(d1[1])) < 1e-9
        assert 0.0 < z1 and np.all(np.isfinite(e1))
    finally:
        h.close()
