"""CPU model of the mixed-precision sweep forms -- TEST INFRASTRUCTURE ONLY.

The reference's sweep (src/3_sweeper/sweeper_simple_c.h:83-277: per octant,
cells in upwind order, m2a, solve, a2m; the solve of
src/3_sweeper/quantities_testing_kernels.h:380-446) evaluated with the
roundings of this library's mixed forms, so that a GPU result can be held to
the arithmetic it claims rather than only to FP64:

  * the solve in the carried form q = r / K, K = S E U the cell's scale (a
    power of two: exact), q = vs/K + q_z/2 + q_y/4 + w(a) q_x, three FP32
    fused multiply-adds in that order; domain-boundary q = (1+a) 2^(a&7);
  * m2a: vs/K = A . (vi/K), the operand split into pieces that each round
    what the pieces before it leave ("bf16": one BF16 piece; "f32+bf16": the
    FP32 rounding and the BF16 rounding of the rest; "f64": no rounding);
    the products are exact and the sum of one 16-angle block is rounded to
    FP32 once;
  * a2m per 16-angle block, the solved values rounded as named ("bf16" or
    "f32"/"f64"), block sums rounded to FP32 and added in FP32;
  * vo summed over octants in FP32 (octant order 0..7), the last octant's
    contribution added in FP64.

With every rounding off (pieces "f64", solve and sums in FP64) the model is
the oracle's algorithm; tests/test_oracle.py pins it to the oracle.  Loops
over cells in Python: for small grids only.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round to BF16 (round to nearest even), through FP32, as float64."""
    f = np.asarray(x, dtype=np.float64).astype(np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def _pieces(x, kind):
    if kind == "f64":
        return [x]
    if kind == "f32":
        return [f32(x)]
    if kind == "bf16":
        return [bf16(x)]
    if kind == "f32+bf16":
        h = f32(x)
        return [h, bf16(x - h)]
    raise ValueError(kind)


def _sf_space(x, y, z):
    return 1 << ((x + 3 * y + 7 * z + 2) & 3)


def _sf_energy(ie):
    return 1 << (((ie * 1366 + 150889) % 714025) & 3)


def _sf_unknown(iu):
    return 1 << (((iu * 741 + 66037) % 312500) & 3)


FORMS = {
    # form: (m2a operand pieces, a2m operand rounding, FP32 arithmetic)
    "fp64": ("f64", "f64", False),
    "fp32": ("f32+bf16", "f32", True),
    "bf16": ("bf16", "bf16", True),
}


def sweep(dims, a_from_m, m_from_a, vi, form="bf16"):
    """vo of one sweep of vi in the given form.  dims: (ncx, ncy, ncz, ne,
    nm, na, nu); a_from_m, m_from_a and vi in the reference layouts (as
    oracle.init_matrices / oracle.init_state give them)."""
    nx, ny, nz, ne, nm, na, nu = dims
    m2a_kind, a2m_kind, is32 = FORMS[form]
    rnd = f32 if is32 else (lambda t: t)
    A = np.asarray(a_from_m, dtype=np.float64).reshape(8, nm, na)   # A[o][m][a]
    M = np.asarray(m_from_a, dtype=np.float64).reshape(8, na, nm)   # M[o][a][m]
    v = np.asarray(vi, dtype=np.float64).reshape(nz, ne, ny, nx, nu, nm)
    eu = (np.array([_sf_energy(i) for i in range(ne)], dtype=np.float64)[:, None] *
          np.array([_sf_unknown(i) for i in range(nu)], dtype=np.float64)[None, :])
    acc = np.zeros((nz, ne, ny, nx, nu, nm))
    vo = np.zeros_like(acc)
    ia = np.arange(na)
    qb = ((1 + ia) * (1 << (ia & 7))).astype(np.float64)
    wz = 0.25 - 2.0 ** (-(ia & 7))
    for o in range(8):
        xs = range(nx - 1, -1, -1) if o & 1 else range(nx)
        ys = range(ny - 1, -1, -1) if o & 2 else range(ny)
        zs = range(nz - 1, -1, -1) if o & 4 else range(nz)
        fz = np.broadcast_to(qb, (ny, nx, ne, nu, na)).copy()
        fy = np.broadcast_to(qb, (nz, nx, ne, nu, na)).copy()
        fx = np.broadcast_to(qb, (nz, ny, ne, nu, na)).copy()
        for z in zs:
            for y in ys:
                for x in xs:
                    k = _sf_space(x, y, z) * eu
                    vc = v[z, :, y, x] / k[:, :, None]
                    vs = rnd(sum(p @ A[o] for p in _pieces(vc, m2a_kind)))
                    t = rnd(vs + 0.5 * fz[y, x])
                    t = rnd(t + 0.25 * fy[z, x])
                    q = rnd(t + wz * fx[z, y])
                    fz[y, x] = q
                    fy[z, x] = q
                    fx[z, y] = q
                    part = np.zeros((ne, nu, nm))
                    for b0 in range(0, na, 16):
                        blk = _pieces(q[..., b0:b0 + 16], a2m_kind)[0] @ M[o, b0:b0 + 16]
                        part = rnd(part + rnd(blk))
                    contrib = part * k[:, :, None]
                    if o < 7:
                        acc[z, :, y, x] = rnd(acc[z, :, y, x] + contrib)
                    else:
                        vo[z, :, y, x] = acc[z, :, y, x] + contrib
    return vo.reshape(-1)
