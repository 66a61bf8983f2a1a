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


def closed_form_normsq(ncx: int, ncy: int, ncz: int, ne: int, nm: int, nu: int) -> int:
    """normsq of the manufactured state (a sweep maps it to itself):
    sum_im (1+im)^2 * sum_cells 4^((x+3y+7z+2)&3) * sum_ie E(ie)^2 * sum_iu U(iu)^2,
    the cell sum by counting residues mod 4 of x, 3y and 7z."""
    def residues(n, mult):
        c = [0, 0, 0, 0]
        for r in range(4):
            # values v in [0, n) with (mult * v) % 4 == r
            c[(mult * r) % 4] += (n - r + 3) // 4 if n > r else 0
        return c
    cx, cy, cz = residues(ncx, 1), residues(ncy, 3), residues(ncz, 7)
    s_x = 0
    for a in range(4):
        for b in range(4):
            for c in range(4):
                s_x += cx[a] * cy[b] * cz[c] * 4 ** ((a + b + c + 2) & 3)
    s_m = sum((1 + im) ** 2 for im in range(nm))
    s_e = sum((1 << (((ie * 1366 + 150889) % 714025) & 3)) ** 2 for ie in range(ne))
    s_u = sum((1 << (((iu * 741 + 66037) % 312500) & 3)) ** 2 for iu in range(nu))
    return s_m * s_x * s_e * s_u
