"""Synthetic problem sets of the batched and served forms.

Uneven problems: per axis ncell = round(exp(U(ln 8, ln 48))), ne uniform in
{8, 16, 32}, na in {32, 64}, nm in {4, 16}, nu = 4; drawn per problem in the
order ncell_x, ncell_y, ncell_z, ne, na, nm from numpy's default_rng(seed).
Arrivals of the served form are a Poisson process (exponential gaps).
"""
from __future__ import annotations

import math

import numpy as np


def uneven_problems(n: int, seed: int):
    """List of (ncell_x, ncell_y, ncell_z, ne, nm, na, nu) tuples."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        c = [int(round(math.exp(rng.uniform(math.log(8), math.log(48))))) for _ in range(3)]
        ne = int(rng.choice([8, 16, 32]))
        na = int(rng.choice([32, 64]))
        nm = int(rng.choice([4, 16]))
        out.append((c[0], c[1], c[2], ne, nm, na, 4))
    return out


def poisson_arrivals(n: int, rate_per_s: float, seed: int):
    """Arrival times (seconds from 0) of n requests at the given mean rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, n))
