"""The learned closure's CPU oracle pinned without running it against
itself: finite-difference gradients, Adam's closed-form first step, the
bfloat16 rounding rule and the box filter."""
import math

import numpy as np
import pytest

from oracle import sgs_mlp as m


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.140625, 1e-40, 65504.0])
    got = m.bf16(x)
    # 1+2^-8 is a tie: rounds to the even neighbour 1.0; 1+3*2^-8 -> 1+2^-6
    assert list(got[:4]) == [1.0, 1.0, 1.0 + 2.0 ** -6, -2.5]
    assert got[4] == 3.140625
    assert got[6] == 65536.0


def test_gradients_match_finite_differences():
    """FP64 mode: every gradient entry (a sample of each matrix) against
    central differences of the loss; catches a transposed operand or a
    missing GELU derivative."""
    rng = np.random.default_rng(3)
    p = m.init_params(1)
    p["b1"] = rng.standard_normal(64) * 0.1
    p["b2"] = rng.standard_normal(64) * 0.1
    p["b3"][:6] = rng.standard_normal(6) * 0.1
    X = rng.standard_normal((7, 10))
    T = rng.standard_normal((7, 6))
    _, g = m.loss_and_grads(p, X, T, exact=True)
    h = 1e-6
    for k, idx in (("W1", [(0, 0), (9, 63), (4, 17)]), ("b1", [(5,), (40,)]),
                   ("W2", [(0, 0), (63, 1), (20, 45)]), ("b2", [(3,), (63,)]),
                   ("W3", [(0, 0), (63, 5), (31, 2)]), ("b3", [(0,), (5,)])):
        for ix in idx:
            q = {kk: vv.copy() for kk, vv in p.items()}
            q[k][ix] += h
            lp, _ = m.loss_and_grads(q, X, T, exact=True)
            q[k][ix] -= 2 * h
            lm, _ = m.loss_and_grads(q, X, T, exact=True)
            fd = (lp - lm) / (2 * h)
            assert abs(g[k][ix] - fd) < 1e-7 + 1e-6 * abs(fd), (k, ix, g[k][ix], fd)
    # the padded rows/columns get no gradient
    assert not g["W1"][10:].any() and not g["W3"][:, 6:].any() and not g["b3"][6:].any()


def test_adam_first_step_closed_form():
    """From zero moments the bias-corrected first step is lr * g / (|g| + eps)."""
    p = {"w": np.array([1.0, -2.0, 0.5])}
    g = {"w": np.array([0.3, -1e-3, 0.0])}
    mm = {"w": np.zeros(3)}
    vv = {"w": np.zeros(3)}
    m.adam(p, g, mm, vv, 1, lr=1e-3)
    want = np.array([1.0, -2.0, 0.5]) - 1e-3 * g["w"] / (np.abs(g["w"]) + 1e-8)
    assert np.allclose(p["w"], want, rtol=0, atol=1e-15)


def test_gelu_known_values():
    assert m.gelu(np.array([0.0]))[0] == 0.0
    assert abs(m.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    assert abs(m.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15


def test_box_filter():
    n, r = 16, 4
    x = np.arange(n) * 2 * np.pi / n
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    u = np.zeros((n, n, n + 2))
    u[:, :, :n] = 2.0 + np.sin(xx)
    f = m.box_filter(u, r)
    assert f.shape == (4, 4, 4)
    assert abs(f.mean() - 2.0) < 1e-14
    # average of sin over cell [x0, x0 + 4 dx) at the grid points
    c0 = np.mean(np.sin(x[:4]))
    assert abs(f[0, 0, 0] - 2.0 - c0) < 1e-14 and abs(f[3, 2, 0] - 2.0 - c0) < 1e-14


def test_backscatter_clip_zeroes_negative_dissipation():
    """After the clip the subgrid dissipation -tau:S is >= 0 everywhere, points
    that dissipated keep their stress, and the clipped stress keeps its part
    orthogonal to S."""
    rng = np.random.default_rng(9)
    g = rng.standard_normal((500, 9))
    tau = rng.standard_normal((500, 6))
    out = m.clip_backscatter(tau, g)
    G = g.reshape(-1, 3, 3)
    S = np.stack([G[:, 0, 0], 0.5 * (G[:, 0, 1] + G[:, 1, 0]), 0.5 * (G[:, 0, 2] + G[:, 2, 0]),
                  G[:, 1, 1], 0.5 * (G[:, 1, 2] + G[:, 2, 1]), G[:, 2, 2]], axis=1)
    w = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])
    d0 = -(w * tau * S).sum(1)
    d1 = -(w * out * S).sum(1)
    assert (d1 >= -1e-12).all()
    assert np.array_equal(out[d0 >= 0], tau[d0 >= 0])
    assert np.abs(d1[d0 < 0]).max() < 1e-12
    assert (d0 < 0).sum() > 100


def test_blend_smagorinsky_limits():
    """The blend with weight 1 is les_model 1's stress -2 nu_t S_ij with the
    reference path's eddy viscosity (turb_visc_smag, m_les.f90:1142-1215),
    and with weight 0 the network's stress itself; catches a missing factor
    2 in |S| or in the symmetric double sum."""
    from oracle import hit3d_oracle as o
    n, cs = 8, 0.17
    rng = np.random.default_rng(5)
    phys = np.zeros((3, n, n, n + 2))
    phys[:, :, :, :n] = rng.standard_normal((3, n, n, n))
    spec = np.stack([o.fft_forward(q) for q in phys])
    spec = o.dealias_all(o.pressure(spec, n), n // 3, n)
    grads = o.velocity_gradients(spec, n)
    g = np.stack([a.ravel() for a in grads], axis=1)
    tv = o.turb_visc_smag(spec, n, cs).ravel()
    tau = rng.standard_normal((n ** 3, 6))
    cd2 = (cs * 2 * np.pi / n) ** 2
    full = m.blend_smagorinsky(tau, g, 1.0, cd2)
    G = g.reshape(-1, 3, 3)
    for c, (i, j) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        S = 0.5 * (G[:, i, j] + G[:, j, i])
        assert np.allclose(full[:, c], -2.0 * tv * S, rtol=1e-12, atol=1e-14)
    assert np.array_equal(m.blend_smagorinsky(tau, g, 0.0, cd2), tau)
    half = m.blend_smagorinsky(tau, g, 0.5, cd2)
    assert np.allclose(half, 0.5 * tau + 0.5 * full, rtol=1e-14, atol=1e-15)
