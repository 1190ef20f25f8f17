"""Pins of the oracle's recurrence search (oracle/or_rec.c) against known
answers and independent expressions:
- the residual table against the norm of the difference exp(-i k.s) Z - Z_j
  computed directly with numpy (the oracle expands it as the reference does,
  with the masked, unmasked and Boussinesq-weighted norms);
- planted relative periodic episodes (tests/rec_cases.py) whose period, shift,
  guess state and residual are known in closed form;
- the reference's lag rules: the newest entry is compared with entries 1 ..
  nStore-1 (never entry 0 before the ring wraps), periods above 20 only at
  even lags, Boussinesq periods below 0.5 skipped.
"""
import numpy as np
import pytest

import oracle_lib as O
from rec_cases import shift_grid, subset_index, wave_history


def run_history(rec, hist, nout, dt):
    recs, checks = [], []
    for n in range(1, len(hist)):
        g, gs, c = rec.update(n * nout, dt, hist[n].reshape(hist.shape[1], -1), want_state=True)
        checks.append(c)
        if g is not None:
            recs.append((g, gs))
    return recs, checks


def direct_residuals(zc, olds, idx, kvec, shifts, nweights, den_fn):
    """res[j, s] = sqrt(normFN / den) with normFN written as |D - S_j|^2_w plus the
    history-norm correction, D = exp(-i k.s) zc on the subset."""
    kx, ky, kz = kvec
    SX, SY, SZ = shifts
    th = np.outer(SX, kx[idx]) + np.outer(SY, ky[idx]) + np.outer(SZ, kz[idx])   # (ns, nnorm)
    ph = np.exp(-1j * th)
    w = nweights["field"][:, None]
    out = np.empty((len(olds), len(SX)))
    for j, zj in enumerate(olds):
        D = ph[:, None, :] * zc[None, :, idx]                      # (ns, nf, nnorm)
        diff = np.sum(w[None] * np.abs(D - zj[None, :, idx]) ** 2, axis=(1, 2))
        corr = nweights["hist_norm"](zj) - np.sum(w * np.abs(zj[:, idx]) ** 2)
        out[j] = np.sqrt((diff + corr) / den_fn(zj))
    return out


def grid3_tables(n):
    g = O.grid3(n, n, n)
    kx, ky, kz, L = O.tables(g)
    return g, (kx, ky, kz), L


def test_ns3d_residual_table_matches_difference_norm():
    g, kvec, L = grid3_tables(24)
    rng = np.random.default_rng(3)
    hist = np.array([(rng.standard_normal((3, g.nkt)) + 1j * rng.standard_normal((3, g.nkt))) * L
                     for _ in range(7)])
    nshift = (30, 2, 2)
    rec = O.Rec(0, g, L, hist[0], store_size=50, norm_width=8, nshift=nshift, nout=10, dt=0.01, threshold=0.1)
    for n in range(1, 7):
        rec.update(n * 10, 0.01, hist[n])
    res, slots = rec.last()
    # the check of update 6 compares entry 5 (pushed by update 5) with entries 4 .. 1
    assert list(slots) == [4, 3, 2, 1]
    idx = subset_index(g, 8)
    hn = lambda z: np.sum(np.abs(z[:, idx]) ** 2 * L[idx][None])   # masked history norm
    ref = direct_residuals(hist[5], hist[list(slots)], idx, kvec, shift_grid(nshift, 1.0),
                           {"field": np.ones(3), "hist_norm": hn}, lambda zj: hn(hist[5]))
    assert np.allclose(res, ref, rtol=1e-12, atol=0)


def test_ns3d_mask_quirk_history_norm_is_masked_only():
    # with energy on masked subset modes the expanded residual keeps |Z_j|^2 masked
    # but |Z_s|^2 and the cross term unmasked (functions.h:111-125 vs 305-320)
    g, kvec, L = grid3_tables(24)
    idx = subset_index(g, 8)
    assert (L[idx] == 0).any()
    rng = np.random.default_rng(4)
    hist = np.array([rng.standard_normal((3, g.nkt)) + 1j * rng.standard_normal((3, g.nkt)) for _ in range(5)])
    nshift = (30, 2, 2)
    rec = O.Rec(0, g, L, hist[0], store_size=50, norm_width=8, nshift=nshift, nout=10, dt=0.01, threshold=0.1)
    for n in range(1, 5):
        rec.update(n * 10, 0.01, hist[n])
    res, slots = rec.last()
    hn = lambda z: np.sum(np.abs(z[:, idx]) ** 2 * L[idx][None])
    ref = direct_residuals(hist[3], hist[list(slots)], idx, kvec, shift_grid(nshift, 1.0),
                           {"field": np.ones(3), "hist_norm": hn}, lambda zj: hn(hist[3]))
    assert np.allclose(res, ref, rtol=1e-10, atol=0)


def test_bous_residual_weights_and_short_period_skip():
    g, kvec, L = grid3_tables(24)
    rng = np.random.default_rng(5)
    hist = np.array([(rng.standard_normal((4, g.nkt)) + 1j * rng.standard_normal((4, g.nkt))) * L
                     for _ in range(10)])
    ri, nshift = 0.7, (64, 1, 1)
    rec = O.Rec(1, g, L, hist[0], store_size=50, norm_width=8, nshift=nshift, nout=10, dt=0.01, threshold=0.1,
                ri=ri)
    for n in range(1, 10):
        rec.update(n * 10, 0.01, hist[n])
    res, slots = rec.last()
    # update 9 compares entry 8 at lags 1..7 (periods 0.1 .. 0.7): those below 0.5 are skipped
    assert list(slots) == [3, 2, 1]
    idx = subset_index(g, 8)
    w = np.array([1.0, 1.0, 1.0, ri])

    def hn(z):   # 3D_Boussinesq/src/functions.h:200-206: ri*x*x + y*y for the density
        s = np.sum(np.abs(z[:3, idx]) ** 2)
        return s + np.sum(ri * z[3, idx].real ** 2 + z[3, idx].imag ** 2)

    ref = direct_residuals(hist[8], hist[list(slots)], idx, kvec, shift_grid(nshift, 1.0),
                           {"field": w, "hist_norm": hn}, lambda zj: hn(zj) + hn(hist[8]))
    assert np.allclose(res, ref, rtol=1e-12, atol=0)


def test_ns2d_residual_table_unmasked():
    g = O.grid2(128, 128)
    kx, ky, kz, L = O.tables(g)
    rng = np.random.default_rng(6)
    hist = np.array([(rng.standard_normal((1, g.nkt)) + 1j * rng.standard_normal((1, g.nkt))) for _ in range(6)])
    nshift = (30, 4, 1)
    rec = O.Rec(2, g, L, hist[0], store_size=50, norm_width=42, nshift=nshift, nout=10, dt=0.01, threshold=0.1)
    for n in range(1, 6):
        rec.update(n * 10, 0.01, hist[n])
    res, slots = rec.last()
    assert list(slots) == [3, 2, 1]
    idx = subset_index(g, 42)
    hn = lambda z: np.sum(np.abs(z[:, idx]) ** 2)
    ref = direct_residuals(hist[4], hist[list(slots)], idx, (kx, ky, kz), shift_grid(nshift, 1.0),
                           {"field": np.ones(1), "hist_norm": hn}, lambda zj: hn(hist[4]))
    assert np.allclose(res, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("alpha", [1.0, 0.5])
def test_planted_episode_gives_one_guess_with_known_period_and_shift(alpha):
    n = 24
    g = O.grid3(int(n / alpha), n, n, alpha)
    kx, ky, kz, L = O.tables(g)
    h, P, sxi = 0.1, 1.2, 7
    hist, _ = wave_history(kx, L, 3, 40, h, P, sxi, 30, alpha=alpha, t_break=2.95, seed=11)
    rec = O.Rec(0, g, L, hist[0], store_size=100, norm_width=8, nshift=(30, 2, 2), nout=10, dt=0.01,
                threshold=1e-2)
    recs, checks = run_history(rec, hist, 10, 0.01)
    assert len(recs) == 1
    gd, gs = recs[0]
    assert gd["index"] == 1
    assert abs(gd["period"] - P) < 1e-12
    assert gd["sx"] == 2 * np.pi * sxi / (30 * alpha) and gd["sy"] == 0.0 and gd["sz"] == 0.0
    # the sequence holds the checks of entries c with c - 12 >= 1 and t_c <= t_break;
    # its best residual is the largest |z(t_c)| one, eps P |N| / |z(t_c)|
    idx = subset_index(g, 8)
    cands = [c for c in range(13, 40) if c * h <= 2.95]
    direct = []
    for c in cands:
        D = np.exp(-1j * kx[idx] * gd["sx"])[None] * hist[c][:, idx]
        direct.append(np.linalg.norm(D - hist[c - 12][:, idx]) /
                      np.sqrt(np.sum(np.abs(hist[c][:, idx]) ** 2 * L[idx][None])))
    best = int(np.argmin(direct))
    cstar = cands[best]
    assert abs(gd["residual"] - direct[best]) < 1e-7 * direct[best]
    assert gd["store"] == cstar - 12
    assert np.array_equal(gs, hist[cstar - 12])
    # the record's time is that of the compared entry plus the period (functions.h:350)
    assert abs(gd["time"] - (cstar * h + P)) < 1e-12
    # the sequence ended at the first check past the break
    end = [i for i, c in enumerate(checks) if c["checked"] and c["time"] > 2.95][0]
    assert checks[end]["residual"] > 1e-2


@pytest.mark.parametrize("P,expect", [(21.0, 42.0), (22.0, 22.0)])
def test_periods_over_20_compared_at_even_lags_only(P, expect):
    g, kvec, L = grid3_tables(24)
    h = 1.0                                    # nout 10, dt 0.1
    hist, _ = wave_history(kvec[0], L, 3, 62, h, P, 4, 30, t_break=55.5, seed=12)
    rec = O.Rec(0, g, L, hist[0], store_size=100, norm_width=8, nshift=(30, 2, 2), nout=10, dt=0.1,
                threshold=1e-2)
    recs, _ = run_history(rec, hist, 10, 0.1)
    assert len(recs) == 1
    assert abs(recs[0][0]["period"] - expect) < 1e-9
    assert recs[0][0]["sx"] == 2 * np.pi * ((4 * round(expect / P)) % 30) / 30


def test_ns2d_planted_episode_and_time_under():
    g = O.grid2(128, 128)
    kx, ky, kz, L = O.tables(g)
    h, P = 0.1, 1.0
    hist, _ = wave_history(kx, L, 1, 40, h, P, 5, 30, t_break=2.55, seed=13)
    rec = O.Rec(2, g, L, hist[0], store_size=100, norm_width=42, nshift=(30, 4, 1), nout=10, dt=0.01,
                threshold=1e-2, v2=0.025)
    recs, checks = run_history(rec, hist, 10, 0.01)
    assert len(recs) == 1
    gd = recs[0][0]
    assert abs(gd["period"] - P) < 1e-12 and gd["sx"] == 2 * np.pi * 5 / 30 and gd["sy"] == 0.0
    assert gd["v2"] == 0.025
    # timeUNDER = |(t_end + c_end h) - (t_start + c_start h)| / period
    # (2D_NavierStokes/src_RK4/functions.h:265-293): the sequence starts at the
    # check of entry 11 (lag 10) and ends at the first check past the break
    start = [c for c in checks if c["checked"] and c["residual"] < 1e-2][0]
    assert abs(start["time"] - 1.1) < 1e-12 and start["count"] == 10
    end = [c for c in checks if c["checked"] and c["time"] > 2.55][0]
    want = abs((end["time"] + end["count"] * h) - (start["time"] + start["count"] * h)) / P
    assert abs(gd["time_under"] - want) < 1e-12


def test_ring_wraps_and_compares_the_oldest_entries():
    # store_size 8: after the ring is full the newest entry's slot rotates and
    # iStart follows it (cuda.cu:120-130); the check compares the 7 other slots
    g, kvec, L = grid3_tables(24)
    rng = np.random.default_rng(7)
    hist = np.array([(rng.standard_normal((3, g.nkt)) + 1j * rng.standard_normal((3, g.nkt))) * L
                     for _ in range(12)])
    rec = O.Rec(0, g, L, hist[0], store_size=8, norm_width=8, nshift=(30, 2, 2), nout=10, dt=0.01, threshold=0.1)
    seen = []
    for n in range(1, 12):
        rec.update(n * 10, 0.01, hist[n])
        seen.append(list(rec.last()[1]))
    # update 8: nStore 7 (entries 1..7 in slots 1..7), compares slot 7 with 6..1
    assert seen[7] == [6, 5, 4, 3, 2, 1]
    # update 9: nStore 8 = store_size: entry 8 went to slot 0, iStart 0; compares slot 0 with 7..1
    assert seen[8] == [7, 6, 5, 4, 3, 2, 1]
    # update 10: entry 9 in slot 1, iStart 1; compares slot 1 with slots 0, 7, .., 2
    assert seen[9] == [0, 7, 6, 5, 4, 3, 2]
