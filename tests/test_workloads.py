"""Problem sets of the batched and served forms (host only)."""
import numpy as np

from minisweep_amd.workloads import poisson_arrivals, uneven_problems


def test_uneven_problems_deterministic_and_in_range():
    a = uneven_problems(256, 0)
    assert a == uneven_problems(256, 0)
    assert a != uneven_problems(256, 1)
    assert len(a) == 256
    for ncx, ncy, ncz, ne, nm, na, nu in a:
        assert all(8 <= c <= 48 for c in (ncx, ncy, ncz))
        assert ne in (8, 16, 32) and na in (32, 64) and nm in (4, 16) and nu == 4
    # log-uniform per axis: the median is near sqrt(8*48) ~ 19.6
    med = np.median([c for s in a for c in s[:3]])
    assert 15 <= med <= 25
    assert {s[4] for s in a} == {4, 16} and {s[5] for s in a} == {32, 64}


def test_poisson_arrivals():
    t = poisson_arrivals(10000, 200.0, 1)
    assert np.all(np.diff(t) > 0)
    assert abs(t[-1] / len(t) - 1 / 200.0) < 0.05 / 200.0 * 10


def test_closed_form_normsq_matches_brute_force():
    from minisweep_amd.workloads import closed_form_normsq
    rng = np.random.default_rng(5)
    for _ in range(100):
        ncx, ncy, ncz = (int(v) for v in rng.integers(1, 14, 3))
        ne, nm, nu = int(rng.integers(1, 6)), int(rng.choice([4, 16])), int(rng.integers(1, 5))
        x, y, z = np.meshgrid(np.arange(ncx), np.arange(ncy), np.arange(ncz), indexing="ij")
        s_x = int(np.sum(4 ** ((x + 3 * y + 7 * z + 2) & 3)))
        s_m = sum((1 + im) ** 2 for im in range(nm))
        s_e = sum((1 << (((ie * 1366 + 150889) % 714025) & 3)) ** 2 for ie in range(ne))
        s_u = sum((1 << (((iu * 741 + 66037) % 312500) & 3)) ** 2 for iu in range(nu))
        assert closed_form_normsq(ncx, ncy, ncz, ne, nm, nu) == s_m * s_x * s_e * s_u
    assert closed_form_normsq(64, 64, 64, 64, 16, 4) == 1541374843289600
