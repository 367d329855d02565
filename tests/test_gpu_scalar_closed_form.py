"""Scalars in a uniform flow through the GPU step, against a closed form:
the velocity is the k = 0 mode alone (no nonlinear term, no viscous decay,
untouched by the projection), and each entry k of a scalar's spectrum
evolves as d(phi_k)/dt = -(i k.u + D |k|^2) phi_k with D = nu/Sc, the
type-0 scalar's mean losing u_1 per unit time (the mean-gradient source,
rhs_scalars.f90:387-389).  Twenty steps (Euler, then AB2) at 64^3 with the
2/3 rule and with the phase shift (the fused scalar passes), and with the
phase shift at 256^3, the size of the two-scalar configuration: each scalar
within 1e-12 of the closed form, the velocity unchanged, the Courant number
|u|_1 dt/dx.  The step size keeps every mode inside AB2's stability region
(D k^2 dt <= 0.35 and |k.u| dt <= 0.35 at 256^3): beyond it the roundoff
of the transforms of the constant products grows from step to step.
"""
import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import FIELDS
from test_oracle_pins import uniform_flow_scalar_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,dealias", [(64, 0), (64, 1), (256, 1)])
def test_uniform_flow_scalars(oracle, n, dealias):
    nu, dt, steps = 0.002, 0.0025, 20
    u, q, sc = (0.3, -0.2, 0.5), (5, 3, -7), (0.7, 2.0)
    print(f"uniform flow scalars n={n} dealias={dealias}", flush=True)
    f0, expected = uniform_flow_scalar_case(oracle, n, u, q, sc, nu)
    h = hit3d_amd.Hit3D(n, n_scalars=2, dealias=dealias)
    try:
        h.set_physics(nu, flow_type=0, pe=tuple(nu / s for s in sc), scalar_type=(0, 1))
        for c in range(5):
            h.put(FIELDS, c + 1, f0[c])
        courant = (abs(u[0]) + abs(u[1]) + abs(u[2])) * dt / (2 * np.pi / n)
        fov = fos = True
        for m in range(1, steps + 1):
            cr, _, fov, fos = h.step(dt, fov, fos, int_scalars=True)
            assert abs(cr - courant) <= 1e-12 * courant, (m, cr)
        ex = expected(steps, dt)
        scale = np.abs(f0[3:]).max()
        for s in range(2):
            err = np.abs(h.get(FIELDS, 4 + s) - ex[s]).max() / scale
            print(f"  scalar {s + 1}: error {err:.2e}", flush=True)
            assert err < 1e-12
        for c in range(3):
            assert np.abs(h.get(FIELDS, c + 1) - f0[c]).max() < 1e-12 * np.abs(f0[:3]).max()
        assert np.abs(h.get(FIELDS, 4) - f0[3]).max() > 1e-2 * scale
    finally:
        h.close()
