"""Synthetic case mixes of the batched and served configurations.

batched_cases  256 cases; n log-uniform in [65, 513] rounded to 2^k + 1 (log2(n-1)
               uniform in [6, 9], rounded), skew A ~ U[0, 0.15], Re ~ U[50, 1000],
               Ri log-uniform in [0.1, 10], Pr in {0.71, 7}; rng seed 4.
served_cases   n log-uniform over the given 2^k + 1 sizes, physics drawn from the
               given ranges (default: the regime where the steady scheme
               converges on these grids, Re ~ U[50, 150], Ri log-uniform in
               [0.1, 1], Pr = 0.71), Poisson arrival times at rate lam; rng seed 5.
p_isolate      a p' problem on one grid: u, v sums of 8 Fourier modes
               sin(kx x + ky y + phi) (kx, ky in 1..8, phi ~ U[0, 2 pi)) that set
               the face weights d, and a N(0, 1) right-hand side minus its mean.
"""
from __future__ import annotations

import numpy as np


def _physics(rng, re=(50.0, 1000.0), ri=(0.1, 10.0), pr=(0.71, 7.0)):
    return dict(Re=float(rng.uniform(*re)),
                Ri=float(np.exp(rng.uniform(np.log(ri[0]), np.log(ri[1])))),
                Pr=float(rng.choice(pr)))


def batched_cases(count=256, seed=4, kmin=6, kmax=9):
    """[(n, skew, physics)] of the batched configuration."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        k = int(np.rint(rng.uniform(kmin, kmax)))
        n = 2 ** k + 1
        skew = float(rng.uniform(0.0, 0.15))
        out.append((n, skew, _physics(rng)))
    return out


def served_cases(count, lam, seed=5, kmin=6, kmax=10, re=(50.0, 150.0), ri=(0.1, 1.0), pr=(0.71,)):
    """[(arrival time s, n, skew, physics)] with Poisson arrivals at rate lam per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / lam))
        k = int(np.rint(rng.uniform(kmin, kmax)))
        n = 2 ** k + 1
        skew = float(rng.uniform(0.0, 0.15))
        out.append((t, n, skew, _physics(rng, re, ri, pr)))
    return out


def p_isolate(grid, seed=20250101):
    """(u, v, q) of the isolated p' problem on `grid`."""
    rng = np.random.default_rng(seed)
    x, y = grid.x, grid.y
    u = np.zeros_like(x)
    v = np.zeros_like(x)
    for f in (u, v):
        for _ in range(8):
            kx, ky = rng.integers(1, 9, size=2)
            ph = rng.uniform(0.0, 2.0 * np.pi)
            f += np.sin(kx * x + ky * y + ph)
    q = rng.standard_normal(x.shape)
    q[1:-1, 1:-1] -= q[1:-1, 1:-1].mean()
    q[0, :] = q[-1, :] = 0.0
    q[:, 0] = q[:, -1] = 0.0
    return u, v, q
