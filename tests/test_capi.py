"""CPU checks of the C ABI: the library loads, exports every entry point the
header declares, and its host-side parts (generator, Gaussian field) match
the oracle bit for bit.  No kernel is launched here."""
import os
import re

import numpy as np
import pytest

import hit3d_amd
from hit3d_amd import lib as hl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RN1 = 67748937.0


def _declared():
    src = open(os.path.join(ROOT, "include", "hit3d.h")).read()
    return sorted(set(re.findall(r"\bint\s+(hit3d_\w+)\s*\(", src)))


def test_library_exports_header_symbols():
    lib = hl.load()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(hl.exported_symbols()) == names


def test_library_exports_measurement_symbols():
    """include/hit3d_measure.h: declared, exported and bound."""
    from hit3d_amd import measure
    src = open(os.path.join(ROOT, "include", "hit3d_measure.h")).read()
    names = sorted(set(re.findall(r"\bint\s+(hit3d_\w+)\s*\(", src)))
    assert names == sorted(measure.exported_symbols()) and names
    lib = hl.load()
    for name in names:
        assert hasattr(lib, name), name


def test_fortran_module_binds_every_entry_point():
    """hit3d_amd/fortran/hit3d_hip.f90 has a bind(C) interface for every
    entry point of both headers."""
    f90 = open(os.path.join(ROOT, "hit3d_amd", "fortran", "hit3d_hip.f90")).read()
    bound = set(re.findall(r'name="(hit3d_\w+)"', f90))
    declared = set(_declared())
    src = open(os.path.join(ROOT, "include", "hit3d_measure.h")).read()
    declared |= set(re.findall(r"\bint\s+(hit3d_\w+)\s*\(", src))
    assert sorted(declared - bound) == []
    assert sorted(bound - declared) == []


def test_uninitialised_call_fails_loudly():
    with pytest.raises(hit3d_amd.Hit3dError, match="hit3d_init"):
        hl.load()
        hl._check(hl._lib.hit3d_pressure())


def test_random_draws_match_oracle(oracle):
    """The product's jump-ahead generator equals the serial restatement, at
    the start and far into the stream (offsets used by a 512^3 fill)."""
    g = oracle.RanduGenerator()
    g.random(-RN1)
    ser = g.draw(RN1, 200000)
    assert np.array_equal(hit3d_amd.random_draws(RN1, 0, 200000), ser)
    assert np.array_equal(hit3d_amd.random_draws(RN1, 123457, 5000), ser[123457:128457])
    g2 = oracle.RanduGenerator()
    g2.random(-RN1)
    g2.draw(RN1, 3 * 10 ** 7)
    far = g2.draw(RN1, 1000)
    assert np.array_equal(hit3d_amd.random_draws(RN1, 3 * 10 ** 7, 1000), far)


def test_random_draws_match_golden_fixture():
    """The library's jump-ahead generator against committed draws of the
    reference's own RANDu.f build (first run of tests/golden/randu_golden.npz:
    index 0 is random(-RN1), index i the i-th draw after it)."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "randu_golden.npz"))
    rn1 = float(g["seeds"][0])
    idx, bits = g["idx1"], g["bits1"]
    head = idx < 4096
    got = hit3d_amd.random_draws(rn1, 0, 4095)
    assert np.array_equal(got.view(np.int64), bits[head][1:])
    for i, b in zip(idx[~head][::7], bits[~head][::7]):
        v = hit3d_amd.random_draws(rn1, int(i) - 1, 1)
        assert v.view(np.int64)[0] == b, i


def test_scalar_gaussian_chain_bit_exact(oracle):
    """The scalar Gaussian fields of init_scalar_spectrum continue the
    generator state left by init_velocity, bit for bit with the oracle."""
    n = 16
    gen = oracle.RanduGenerator()
    oracle.init_velocity_uniforms(n, RN1, gen)
    iy = hl.velocity_rng_state(n, RN1)
    assert iy == gen.iy
    for _ in range(2):
        gen.random(-93712139.0)
        w = gen.draw(93712139.0, 2 * (n + 2) * n * n).reshape(2, n, n, n + 2)
        want = oracle.box_muller(w[0], w[1])
        got, iy = hl.scalar_gaussian(n, 93712139.0, iy)
        assert np.array_equal(got, want)
        assert iy == gen.iy


def test_gaussian_field_bit_exact(oracle):
    """init_velocity.f90:87-113 on one rank: the product's Gaussian field for
    each component equals the oracle's (same draws, glibc Box-Muller)."""
    n = 16
    uni, _ = oracle.init_velocity_uniforms(n, RN1)
    g = oracle.box_muller(uni[0:3], uni[3:6])
    for c in range(3):
        mine = hl.init_gaussian(n, RN1, c + 1)
        assert np.array_equal(mine.view(np.int64), g[c].view(np.int64))
