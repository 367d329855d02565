"""Writes randu_golden.npz: draws of the reference's own generator
(RANDu.f compiled from the reference sources by oracle/Makefile into
oracle/_ref/randu_ref) as 64-bit patterns, for the two seeds the velocity
and scalar initialisations use (00_example.in:32-33), in the order the
reference calls them: random(-RN1), COUNT x random(RN1), random(-RN2),
COUNT x random(RN2).

Stored: the first HEAD values of each run and every STRIDE-th value after
that (index arrays alongside), so the fixture stays small.

randu_chain16.npz: every value of the chain of calls a 16^3 run with two
spectrum-initialised scalars makes (velocity, then two scalars on RN2).
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "randu_ref")
RN1, RN2 = 67748937.0, 93712139.0
COUNT1, COUNT2 = 3_000_000, 500_000
HEAD, STRIDE = 4096, 997


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/randu_ref is missing: run `make -C oracle` where the reference exists")
    out = subprocess.run([REF, repr(RN1), str(COUNT1), repr(RN2), str(COUNT2)], check=True,
                         capture_output=True, text=True).stdout.split()
    bits = np.array([int(v) for v in out], dtype=np.int64)
    a = bits[:1 + COUNT1]
    b = bits[1 + COUNT1:]
    assert len(b) == 1 + COUNT2

    def pick(run):
        idx = np.concatenate([np.arange(HEAD), np.arange(HEAD, len(run), STRIDE)])
        return idx.astype(np.int64), run[idx]

    i1, v1 = pick(a)
    i2, v2 = pick(b)
    np.savez_compressed(os.path.join(HERE, "randu_golden.npz"), seeds=np.array([RN1, RN2]),
                        counts=np.array([COUNT1, COUNT2]), idx1=i1, bits1=v1, idx2=i2, bits2=v2)

    # the chain a 16^3 run with two spectrum scalars makes: init_velocity
    # (random(-RN1), 6 (n+2) n^2 draws), then per scalar random(-RN2) and
    # 2 (n+2) n^2 draws (init_scalars.f90:105-125), all values kept
    per = 18 * 16 * 16
    out = subprocess.run([REF, repr(RN1), str(6 * per), repr(RN2), str(2 * per), repr(RN2),
                          str(2 * per)], check=True, capture_output=True, text=True).stdout.split()
    bits = np.array([int(v) for v in out], dtype=np.int64)
    assert len(bits) == 3 + 10 * per
    np.savez_compressed(os.path.join(HERE, "randu_chain16.npz"), seeds=np.array([RN1, RN2]),
                        n=np.array([16]), bits=bits)


if __name__ == "__main__":
    main()
