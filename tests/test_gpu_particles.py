"""GPU parity of Lagrangian particle tracking (m_particles.f90) through the C
ABI against the oracle on the same physical velocity and positions.

The kernel evaluates the reference's expressions in the same order without
contraction, so positions and velocities are compared bit for bit; the
physical velocity from fields_to_parts is an inverse FFT and is compared to
1e-13 relative.
"""
import sys

import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS, WRK

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_velocity(n, seed):
    rng = np.random.default_rng(seed)
    out = np.zeros((3, n, n, n + 2))
    out[:, :, :, :n] = rng.uniform(-1, 1, (3, n, n, n))
    return out


def random_positions(n, count, seed):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(1.0, n + 1.0, (count, 3))
    # particles on the seams and cell faces
    edge = np.array([[1.0, 1.0, 1.0], [n + 0.999999, n + 0.5, 1.0], [n, n, n],
                     [1.0 + 1e-12, n + 1.0 - 1e-12, 2.0], [n / 2, n / 2 + 0.5, n + 0.25]])
    xyz[:len(edge)] = edge
    return xyz


@pytest.mark.parametrize("n,scheme", [(16, 0), (16, 1), (32, 1), (64, 0)])
def test_particles_move_bit_exact(oracle, n, scheme):
    note(f"particles n={n} scheme={scheme}")
    vel = random_velocity(n, 300 + n)
    count = 5000
    xyz = random_positions(n, count, 400 + n)
    dx = 2 * np.pi / n
    dt = 0.4 * dx   # |1.5u - 0.5u_old| <= 2: moves under one cell per step
    h = hit3d_amd.Hit3D(n)
    try:
        for c in range(3):
            h.put(WRK, c + 1, vel[c])
        h.particles_set(xyz)
        pos, uvw = xyz.copy(), np.full((count, 3), -999.0)
        for step in range(4):   # Euler, then AB2
            h.particles_move(dt, scheme)
            pos, uvw = oracle.particles_move(vel, pos, uvw, dt, n, scheme)
            gx, gu = h.particles_get()
            assert np.array_equal(gu, uvw), f"velocity step {step}"
            assert np.array_equal(gx, pos), f"position step {step}"
            assert (gx >= 1.0).all() and (gx < n + 1.0).all()
    finally:
        h.close()


def test_fields_to_parts_and_move(oracle):
    """fields_to_parts leaves the physical velocity of fields(1:3) in wrk(1:3);
    a move from it matches the oracle's move from the oracle's transform."""
    n = 32
    rng = np.random.default_rng(77)
    phys = np.zeros((3, n, n, n + 2))
    phys[:, :, :, :n] = rng.uniform(-1, 1, (3, n, n, n))
    spec = np.stack([oracle.fft_forward(p) for p in phys])
    h = hit3d_amd.Hit3D(n)
    try:
        for c in range(3):
            h.put(FIELDS, c + 1, spec[c])
        h.fields_to_parts()
        for c in range(3):
            got = h.get(WRK, c + 1)[:, :, :n]
            want = oracle.fft_inverse(spec[c])[:, :, :n]
            assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
        xyz = random_positions(n, 1000, 5)
        h.particles_set(xyz)
        h.particles_move(0.01, 1)
        gx, gu = h.particles_get()
        vel = [oracle.fft_inverse(spec[c]) for c in range(3)]
        wx, wu = oracle.particles_move(vel, xyz, np.full((1000, 3), -999.0), 0.01, n, 1)
        assert np.abs(gu - wu).max() <= 1e-13 * np.abs(wu).max()
        assert np.abs(gx - wx).max() <= 1e-13 * n
    finally:
        h.close()


def test_particles_set_rejects_out_of_range():
    h = hit3d_amd.Hit3D(16)
    try:
        with pytest.raises(hit3d_amd.Hit3dError):
            h.particles_set(np.array([[0.5, 2.0, 2.0]]))
        with pytest.raises(hit3d_amd.Hit3dError):
            h.particles_set(np.array([[2.0, 17.0, 2.0]]))
    finally:
        h.close()
