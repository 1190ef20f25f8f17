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


def test_upozk_records_layout_and_round_trip(tmp_path):
    # UPO_ZK.out is a C stream: 7 doubles then the state, per record
    # (2D_NavierStokes/src_RK4/functions.h:294-301)
    g = O.grid2(32, 32)
    kx, ky, _, L = O.tables(g)
    snaps = np.array([O.ic_ns2d_cfg(g, 40 + i) * L for i in range(4)])
    hits = [{"i": 2, "t": 11.5, "T": 6.25, "shift": 0.3, "dist": 0.12},
            {"i": 0, "t": 3.0, "T": 9.0, "shift": -1.1, "dist": 0.2}]
    p = str(tmp_path / "UPO_ZK.out")
    R.write_upozk(p, hits, snaps, 1 / 40.0)
    raw = np.fromfile(p, np.float64)
    assert raw.size == 2 * (7 + 2 * g.nkt)
    assert list(raw[:7]) == [11.5, 1 / 40.0, 6.25, 0.3, 0.0, 0.12, 0.0]
    assert np.array_equal(raw[7:7 + 2 * g.nkt].view(np.complex128), snaps[2])
    R.write_upozk(p, hits[:1], snaps, 1 / 40.0, append=True)
    recs = R.read_upozk(p, g.nkt)
    assert len(recs) == 3
    assert recs[1][0][2] == 9.0 and np.array_equal(recs[1][1][0], snaps[0])
    assert np.array_equal(recs[2][1][0], snaps[2])
    # a 3D file (three fields per record) read with the wrong field count is refused
    try:
        R.read_upozk(p, g.nkt, nfields=3)
        raise AssertionError("expected ValueError")
    except ValueError:
        pass
