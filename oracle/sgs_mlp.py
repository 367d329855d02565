"""ORACLE -- test infrastructure only.

CPU restatement of the learned subgrid-stress closure (no counterpart in the
reference; it stands in for the eddy viscosity of m_les.f90:1142-1215): a
pointwise MLP 10 -> 64 -> 64 -> 6 with GELU, trained with Adam on standardised
box-filtered DNS data.

Two modes:
* exact=True: FP64 throughout (pinned by finite differences in the tests);
* exact=False: the rounding points of the GPU kernels -- weights, inputs and
  the activations/gradients fed to the matrix units rounded to bfloat16
  (round to nearest even), everything else FP64 -- so the GPU (FP32
  accumulation) can be compared with a tolerance set by FP32 rounding.

Only tests/ may import this module, and only as the checker.
"""
from __future__ import annotations

import math

import numpy as np

NIN, NH, NOUT = 10, 64, 6
KIN, KOUT = 16, 16  # padded widths the matrix units see


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    out = (((u + r) >> 16) << 16).astype(np.uint32)
    return out.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.vectorize(math.erf)(z / math.sqrt(2.0)))


def gelu_grad(z):
    cdf = 0.5 * (1.0 + np.vectorize(math.erf)(z / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)
    return cdf + z * pdf


def init_params(seed: int = 0) -> dict:
    """He-style initialisation (the kernels' layout: W1[16,64] with input rows
    10..15 zero, W3[64,16] with output columns 6..15 zero)."""
    rng = np.random.default_rng(seed)
    p = {"W1": np.zeros((KIN, NH)), "b1": np.zeros(NH), "W2": rng.standard_normal((NH, NH)) * math.sqrt(2.0 / NH),
         "b2": np.zeros(NH), "W3": np.zeros((NH, KOUT)), "b3": np.zeros(KOUT)}
    p["W1"][:NIN] = rng.standard_normal((NIN, NH)) * math.sqrt(2.0 / NIN)
    p["W3"][:, :NOUT] = rng.standard_normal((NH, NOUT)) * math.sqrt(1.0 / NH)
    return p


def _pad_in(X):
    out = np.zeros((X.shape[0], KIN))
    out[:, :NIN] = X
    return out


def forward(p: dict, X: np.ndarray, exact: bool = False):
    """Y[B, 6] and the cache for backward.  X[B, 10] standardised inputs."""
    r = (lambda a: a) if exact else bf16
    Xb = r(_pad_in(X))
    W1, W2, W3 = r(p["W1"]), r(p["W2"]), r(p["W3"])
    Z1 = Xb @ W1 + p["b1"]
    H1 = r(gelu(Z1))
    Z2 = H1 @ W2 + p["b2"]
    H2 = r(gelu(Z2))
    Y = H2 @ W3 + p["b3"]
    return Y[:, :NOUT], dict(Xb=Xb, Z1=Z1, H1=H1, Z2=Z2, H2=H2, W1=W1, W2=W2, W3=W3)


def loss_and_grads(p: dict, X: np.ndarray, T: np.ndarray, exact: bool = False, nb: int | None = None):
    """MSE over the B x 6 outputs (divided by nb*6, nb = B unless given) and
    its gradients with the kernels' rounding points."""
    r = (lambda a: a) if exact else bf16
    B = X.shape[0]
    nb = B if nb is None else nb
    Y, c = forward(p, X, exact)
    D = Y - T
    loss = float(np.sum(D * D)) / (nb * NOUT)
    dY = np.zeros((B, KOUT))
    dY[:, :NOUT] = 2.0 * D / (nb * NOUT)
    dYb = r(dY)
    g = {"W3": c["H2"].T @ dYb, "b3": dY.sum(0)}
    dZ2 = (dYb @ c["W3"].T) * gelu_grad(c["Z2"])
    dZ2b = r(dZ2)
    g["W2"] = c["H1"].T @ dZ2b
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2b @ c["W2"].T) * gelu_grad(c["Z1"])
    dZ1b = r(dZ1)
    g["W1"] = c["Xb"].T @ dZ1b
    g["b1"] = dZ1.sum(0)
    return loss, g


def adam(p: dict, g: dict, m: dict, v: dict, t: int, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, with bias correction), in place."""
    for k in p:
        m[k] = b1 * m[k] + (1.0 - b1) * g[k]
        v[k] = b2 * v[k] + (1.0 - b2) * g[k] * g[k]
        mh = m[k] / (1.0 - b1 ** t)
        vh = v[k] / (1.0 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)


def box_filter(u: np.ndarray, ratio: int):
    """Top-hat average over ratio^3 blocks of a physical field a[z, y, x]
    (x < n), giving the coarse grid n/ratio."""
    n = u.shape[0]
    m = n // ratio
    a = np.asarray(u)[:, :, :n].reshape(m, ratio, m, ratio, m, ratio)
    return a.mean(axis=(1, 3, 5))


def blend_smagorinsky(tau: np.ndarray, grads: np.ndarray, mix: float, cd2: float) -> np.ndarray:
    """(1 - mix) tau + mix tau_S with the Smagorinsky stress tau_S = -2 cd2
    |S| S, |S| = sqrt(2 S:S), cd2 = (c_s dx)^2 (les_model 1's stress,
    m_les.f90:1142-1215 with rhs_velocity's -2 nu_t S_ij); the build's blend of
    the learned closure, which has no reference counterpart.  tau[p, 6] in the
    order 11, 12, 13, 22, 23, 33."""
    g = grads.reshape(-1, 3, 3)
    S = np.stack([g[:, 0, 0], 0.5 * (g[:, 0, 1] + g[:, 1, 0]), 0.5 * (g[:, 0, 2] + g[:, 2, 0]),
                  g[:, 1, 1], 0.5 * (g[:, 1, 2] + g[:, 2, 1]), g[:, 2, 2]], axis=1)
    w = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])
    smag = np.sqrt(2.0 * (w * S * S).sum(1))
    return (1.0 - mix) * tau + mix * (-2.0 * cd2 * smag)[:, None] * S


def clip_backscatter(tau: np.ndarray, grads: np.ndarray) -> np.ndarray:
    """Where the subgrid dissipation -tau:S is negative (S the strain rate
    from grads[p, 3i+j] = du_i/dx_j), add (-tau:S)/(S:S) S: the dissipation
    becomes zero and the part of tau orthogonal to S is kept.  tau[p, 6] in
    the order 11, 12, 13, 22, 23, 33."""
    g = grads.reshape(-1, 3, 3)
    S = np.stack([g[:, 0, 0], 0.5 * (g[:, 0, 1] + g[:, 1, 0]), 0.5 * (g[:, 0, 2] + g[:, 2, 0]),
                  g[:, 1, 1], 0.5 * (g[:, 1, 2] + g[:, 2, 1]), g[:, 2, 2]], axis=1)
    w = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])
    ts = (w * tau * S).sum(1)
    ss = (w * S * S).sum(1)
    clip = (ts > 0) & (ss > 0)
    a = np.where(clip, -ts / np.where(ss > 0, ss, 1.0), 0.0)
    return tau + a[:, None] * S
