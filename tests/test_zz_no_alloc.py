"""The allocation-free entry point hit3d_mlp_grads_static against the
allocating hit3d_mlp_grads, through the C ABI: the same kernels run on a
static scratch array, so gradients and loss agree bit for bit -- on two
calls in a row, the second with fewer samples (the array keeps the first
call's contents) -- and a count above the limit is rejected."""
import ctypes

import numpy as np
import pytest

import hit3d_amd

pytestmark = pytest.mark.gpu


def test_mlp_grads_static_matches_allocating_entry():
    print("mlp_grads_static vs mlp_grads", flush=True)
    h = hit3d_amd.Hit3D(16)
    try:
        rng = np.random.default_rng(11)
        p = (rng.standard_normal(h.NPARAM) * 0.2).astype(np.float32)
        h.mlp_set_params(p)
        for count in (8192, 96):
            X = rng.standard_normal((count, 10)).astype(np.float32)
            T = rng.standard_normal((count, 6)).astype(np.float32)
            g_ref, l_ref = h.mlp_grads(X, T)
            g_st, l_st = h.mlp_grads_static(X, T)
            assert np.array_equal(g_st, g_ref), count
            assert l_st == l_ref, count
            assert np.abs(g_ref).max() > 0.0 and l_ref > 0.0
        X = np.zeros((8224, 10), np.float32)
        T = np.zeros((8224, 6), np.float32)
        with pytest.raises(hit3d_amd.Hit3dError, match="static scratch limit"):
            h.mlp_grads_static(X, T)
        with pytest.raises(hit3d_amd.Hit3dError, match="multiple of 32"):
            h.mlp_grads_static(X[:40], T[:40])
    finally:
        h.close()
