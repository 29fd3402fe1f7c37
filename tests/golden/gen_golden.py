"""Writes the golden vectors the CPU tests pin the oracle against.

Run from the repository root: python tests/golden/gen_golden.py
Every value comes from a library routine or a closed form, not from the oracle:
  natural_spline.npz  scipy.interpolate.CubicSpline(bc_type="natural") values and
                      slopes on a non-uniform grid (einspline's NATURAL B-spline
                      interpolant is the same function)
  kspace_counts.json  brute-force integer counts of half-space k-vectors per |n|^2 shell
  free_images.npz     direct image sums -log1p(sum_n e^{-((x+nL)^2 - x^2)/4 lambda tau})
"""
import json
import math
import os

import numpy as np
from scipy.interpolate import CubicSpline

HERE = os.path.dirname(os.path.abspath(__file__))


def natural_spline():
    rng = np.random.default_rng(20261202)
    x = np.sort(rng.uniform(0.0, 10.0, 60))
    x[0], x[-1] = 0.0, 10.0
    y = np.sin(1.3 * x) * np.exp(-0.1 * x) + 0.05 * rng.normal(size=x.size)
    t = np.concatenate([rng.uniform(0.0, 10.0, 400), x[1:-1], [0.0, 10.0]])
    cs = CubicSpline(x, y, bc_type="natural")
    np.savez(os.path.join(HERE, "natural_spline.npz"), x=x, y=y, t=t, v=cs(t), dv=cs(t, 1))


def kspace_counts():
    out = {}
    for shell in (25, 61, 97, 153):
        m = int(math.isqrt(shell)) + 1
        r = np.arange(-m, m + 1)
        a, b, c = np.meshgrid(r, r, r, indexing="ij")
        n2 = a * a + b * b + c * c
        half = (a > 0) | ((a == 0) & (b > 0)) | ((a == 0) & (b == 0) & (c > 0))
        out[str(shell)] = int(np.count_nonzero((n2 <= shell) & (n2 > 0) & half))
    with open(os.path.join(HERE, "kspace_counts.json"), "w") as f:
        json.dump(out, f, indent=1)


def free_images():
    L, lam, tau, n_images = 3.0, 0.5, 2.0, 4
    x = np.linspace(-L / 2, L / 2, 37)
    s = np.zeros_like(x)
    for n in range(1, n_images + 1):
        for sgn in (1, -1):
            xp = x + sgn * n * L
            s += np.exp((x * x - xp * xp) / (4 * lam * tau))
    np.savez(os.path.join(HERE, "free_images.npz"), L=L, lam=lam, tau=tau, n_images=n_images, x=x, image=-np.log1p(s))


if __name__ == "__main__":
    natural_spline()
    kspace_counts()
    free_images()
    print("golden vectors written to", HERE)
