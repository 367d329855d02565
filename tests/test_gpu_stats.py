"""GPU parity of the diagnostics: scalar statistics (m_stats.f90:295-437) and
the divergence check (pressure.f90:53-93), through the C ABI, against the
oracle on the same seeded fields.

Tolerances: sums and moments 1e-12 relative (the GPU sums per z-plane, then
over planes, in FP64; the oracle sums with numpy's pairwise order), min/max
1e-13 of the field's largest magnitude, spectra 1e-12 relative.
"""
import sys

import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_spec(oracle, n, seed, count):
    rng = np.random.default_rng(seed)
    out = np.zeros((count, n, n, n + 2))
    for c in range(count):
        a = np.zeros((n, n, n + 2))
        a[:, :, :n] = rng.standard_normal((n, n, n)) + 0.3 * c
        out[c] = oracle.fft_forward(a)
    return out


@pytest.mark.parametrize("n,dealias", [(16, 0), (32, 1), (64, 0)])
def test_stat_scalars_match_oracle(oracle, n, dealias):
    note(f"stat_scalars n={n} dealias={dealias}")
    ns = 2
    pe = (0.013, 0.0021)
    f = random_spec(oracle, n, 40 + n, 3 + ns)
    h = hit3d_amd.Hit3D(n, n_scalars=ns, dealias=dealias)
    try:
        h.set_physics(0.01, pe=pe, scalar_type=(0, 0))
        for c in range(3 + ns):
            h.put(FIELDS, c + 1, f[c])
        got = h.stat_scalars()
        kmax = h.kmax
        for s in range(ns):
            want = oracle.scalar_stats(f[3 + s], n, kmax, pe[s])
            for key in ("sc_diss", "mean", "variance"):
                assert abs(got[key][s] - want[key]) <= 1e-12 * abs(want[key]), (key, s)
            scale = max(abs(want["sc_min"]), abs(want["sc_max"]))
            assert abs(got["sc_min"][s] - want["sc_min"]) <= 1e-13 * scale
            assert abs(got["sc_max"][s] - want["sc_max"]) <= 1e-13 * scale
            d = np.abs(got["es_spec"][s] - want["es_spec"]).max()
            assert d <= 1e-12 * np.abs(want["es_spec"]).max()
    finally:
        h.close()


def test_stat_scalars_closed_form():
    """phi = c + a cos x + b sin 2y on the GPU alone: the same closed forms
    the oracle is pinned with."""
    n, a, b, c, pe = 32, 0.7, 0.3, 1.25, 0.02
    x = np.arange(n) * 2 * np.pi / n
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    phys = np.zeros((n, n, n + 2))
    phys[:, :, :n] = c + a * np.cos(xx) + b * np.sin(2 * yy)
    h = hit3d_amd.Hit3D(n, n_scalars=1)
    try:
        h.set_physics(0.01, pe=(pe,), scalar_type=(0,))
        h.put(FIELDS, 4, phys)
        h.xFFT3d_fields(1, 4)
        st = h.stat_scalars()
        assert abs(st["mean"][0] - c) < 1e-13
        assert abs(st["variance"][0] - (a * a + b * b) / 2) < 1e-13
        assert abs(st["es_spec"][0, 0] - a * a / 2) < 1e-14
        assert abs(st["es_spec"][0, 1] - b * b / 2) < 1e-14
        assert abs(st["sc_diss"][0] - 2 * pe * (a * a / 2 + 2 * b * b)) < 1e-14
        assert abs(st["sc_min"][0] - (c - a - b)) < 1e-13
        assert abs(st["sc_max"][0] - (c + a + b)) < 1e-13
    finally:
        h.close()


@pytest.mark.parametrize("n", [16, 64])
def test_divergence_matches_oracle(oracle, n):
    note(f"divergence n={n}")
    f = random_spec(oracle, n, 90 + n, 3)
    h = hit3d_amd.Hit3D(n)
    try:
        for c in range(3):
            h.put(FIELDS, c + 1, f[c])
        lo, hi = h.divergence()
        wlo, whi = oracle.divergence(f, n)
        scale = max(abs(wlo), abs(whi))
        assert abs(lo - wlo) <= 1e-13 * scale and abs(hi - whi) <= 1e-13 * scale
        # after the projection the divergence is round-off
        h.pressure()
        lo, hi = h.divergence()
        assert max(abs(lo), abs(hi)) < 1e-11 * scale
    finally:
        h.close()
