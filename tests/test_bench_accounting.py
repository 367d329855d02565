"""The byte accounting behind bench.py's roofline figures, pinned by hand
counts (no GPU): algorithmic bytes per launch of each fused-step kernel
class at 512^3 (K = 170, 171 kept kx columns, 341 kept ky/kz rows), the
reference's unpruned byte model, and the refusal of more than one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

N, NK, NKY = 512, 171, 341


def test_kernel_bytes_512_deviatoric():
    cube = NK * NKY * NKY * 16           # one FP64 field inside the 2/3 cube
    zy = N * NKY * NK * 16               # a work buffer with kept ky rows, all z
    full = N * N * NK * 16               # a work buffer, all y and z
    want = {
        "step_inv_z": 3 * cube + 3 * zy,
        "step_inv_y": 3 * zy + 3 * full,
        "step_x": 3 * full + 5 * full,
        "step_fwd_y": 5 * full + 5 * zy,
        "step_fwd_z": 5 * zy + 6 * cube + 6 * cube,
    }
    for cls, b in want.items():
        assert bench.kernel_bytes(N, cls, 0, 2) == b, cls
    # the x pass's 5.74 GB and the step's 23.9 GB quoted with the roofline
    assert bench.kernel_bytes(N, "step_x", 0, 2) == 5737807872
    assert abs(sum(want.values()) - 23.89e9) < 0.01e9


def test_kernel_bytes_mixed_and_forms():
    # mixed: FP32 work buffers halve every term but the fields; rhs_old is FP32
    cube64 = NK * NKY * NKY * 16
    cube32 = NK * NKY * NKY * 8
    zy32 = N * NKY * NK * 8
    assert bench.kernel_bytes(N, "step_fwd_z", 1, 2) == 5 * zy32 + 6 * cube64 + 6 * cube32
    # divergence form: six products; rotational: six inverse, three forward
    full = N * N * NK * 16
    assert bench.kernel_bytes(N, "step_x", 0, 0) == 9 * full
    assert bench.kernel_bytes(N, "step_x", 0, 1) == 9 * full


def test_unpruned_model_and_slot():
    s = bench.slot_bytes(N)
    assert s == 514 * 512 * 512 * 8
    assert 6 * 9 * s + 4 * 3 * s == 71143784448   # the measurement plan's 66 S at 512^3


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr
