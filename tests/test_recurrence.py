"""CPU tests of the recurrence search (psgpu_amd/recurrence.py) on synthetic
states with a known shift and period."""
import numpy as np

import oracle_lib as O
from psgpu_amd import recurrence as R


def test_shift_recovered_exactly():
    g = O.grid2(64, 64)
    kx, ky, _, L = O.tables(g)
    a = O.ic_ns2d_cfg(g, 3) * L
    s0 = 2 * np.pi * 37 / 256
    b = a * np.exp(1j * kx * s0)            # S_{s0} b = a
    ov, sh = R.shifted_overlap(a, b[None], g.iktx, L, M=256)
    assert abs(sh[0] - s0) < 1e-12
    assert abs(ov[0] - np.sum(np.abs(a) ** 2 * L)) < 1e-9 * ov[0]


def test_best_recurrence_finds_planted_period():
    g = O.grid2(32, 32)
    kx, ky, _, L = O.tables(g)
    rng = np.random.default_rng(5)
    times = np.arange(0, 30.0, 0.5)
    snaps = np.array([O.ic_ns2d_cfg(g, 100 + i) * L for i in range(len(times))])
    s0 = 2 * np.pi * 11 / 512
    i0, T0 = 6, 8.5
    j0 = int(np.nonzero(np.isclose(times, times[i0] + T0))[0][0])
    snaps[j0] = snaps[i0] * np.exp(1j * kx * s0) + 1e-6 * rng.standard_normal(g.nkt) * L
    hits = R.best_recurrences(times, snaps, g.iktx, L, 5.0, 20.0, top=3)
    h = hits[0]
    assert h["i"] == i0 and abs(h["T"] - T0) < 1e-12 and abs(h["shift"] - s0) < 1e-12 and h["dist"] < 1e-4
    assert all(o["dist"] > 0.5 for o in hits[1:])
