"""Writes tests/golden/bernoulli.npz: high-precision B(x) = x/expm1(x) and B'(x).

Restates the goldens of test/test010_bernoulli.jl:6-21 (BigFloat B on
-1:1.00001e-5:1 and -100:1.00001e-3:100) with mpmath at 60 digits, every 100th
point of each range, plus the derivative formula of
pluto-examples/bernoulli.jl:45-51, DB(x) = -(x e^x - e^x + 1)/(e^x - 1)^2.
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 60


def B(x):
    bx = mp.mpf(float(x))
    if bx == 0:
        return 1.0
    return float(bx / mp.expm1(bx))


def DB(x):
    bx = mp.mpf(float(x))
    if bx == 0:
        return -0.5
    e = mp.exp(bx)
    return float(-(bx * e - e + 1) / ((e - 1) * (e - 1)))


def main():
    small = -1.0 + 1.00001e-5 * np.arange(0, 200000, 100)
    large = -100.0 + 1.00001e-3 * np.arange(0, 200000, 100)
    x = np.concatenate([small, large, [0.0, 0.2499999, 0.25, -0.25, 49.999, 50.0, 50.001, -50.001]])
    b = np.array([B(v) for v in x])
    db = np.array([DB(v) for v in x])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bernoulli.npz")
    np.savez_compressed(out, x=x, b=b, db=db)


if __name__ == "__main__":
    main()
