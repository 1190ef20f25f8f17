"""GPU recurrence search (include/psg_rec.h) against the oracle (oracle/or_rec.c):
the residual table of every check, the check summaries, the guess records and
their states, on planted relative periodic episodes (tests/rec_cases.py) and
on DNS histories stepped on the GPU.  Residuals are compared as res^2 =
normFN / |Z|^2, an O(1) quantity whose rounding error is absolute (1e-13);
record fields that are integers, shifts taken from the grid and guess states
must be equal."""
import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P
from rec_cases import wave_history

pytestmark = pytest.mark.gpu

RULE = {P.NS3D: 0, P.EULER3D: 0, P.BOUS3D: 1, P.NS2D: 2}


def make(flavour, n, dt=0.01, v2=0.02, ri=0.0):
    kw = dict(flavour=flavour, nx=n, ny=n, nz=1 if flavour == P.NS2D else n, dt=dt, v2=v2)
    if flavour == P.BOUS3D:
        kw.update(ri=ri, sc=1.0, theta=np.pi / 2)
    return P.Stepper(**kw)


def ogrid(flavour, n):
    return O.grid2(n, n) if flavour == P.NS2D else O.grid3(n, n, n)


def compare_run(st, g, L, hist_fn, nupdates, nout, dt, opts, ri=0.0):
    """Drive the GPU search and the oracle over the same entries; hist_fn(e) gives
    entry e after the stepper is at it (it may step the GPU)."""
    s0 = hist_fn(0)
    st.set_state(s0)
    rec = P.Recurrence(st, nout=nout, **opts)
    o = rec.opts
    orr = O.Rec(o.rule, g, L, s0, store_size=o.store_size, norm_width=o.norm_width,
                nshift=(o.nshift_x, o.nshift_y, o.nshift_z), nout=nout, dt=dt, tstart=o.tstart,
                threshold=o.threshold, ri=ri, v2=st.cfg.v2)
    worst, nrec, nchk = 0.0, 0, 0
    for e in range(1, nupdates):
        s = hist_fn(e)
        gg, gs, gc = rec.update(e * nout, dt, want_state=True)
        og, ogs, oc = orr.update(e * nout, dt, s, want_state=True)
        assert (gg is None) == (og is None), e
        for k in ("checked", "compared", "store", "count"):
            assert gc[k] == oc[k], (e, k, gc, oc)
        if gc["checked"] and gc["compared"]:
            nchk += 1
            assert gc["sx"] == oc["sx"] and gc["sy"] == oc["sy"] and gc["sz"] == oc["sz"]
            assert abs(gc["residual"] ** 2 - oc["residual"] ** 2) < 1e-13
            assert abs(gc["time"] - oc["time"]) < 1e-12
        res, slots = rec.last()
        ores, oslots = orr.last()
        assert np.array_equal(slots, oslots)
        if res.size:
            worst = max(worst, float(np.max(np.abs(res ** 2 - ores ** 2))))
        if gg is not None:
            nrec += 1
            for k in ("index", "store"):
                assert gg[k] == og[k]
            for k in ("sx", "sy", "sz"):
                assert gg[k] == og[k]
            for k in ("time", "v2", "period", "time_under"):
                assert abs(gg[k] - og[k]) < 1e-12 * max(1.0, abs(og[k])), k
            assert abs(gg["residual"] ** 2 - og["residual"] ** 2) < 1e-13
            assert np.array_equal(gs, ogs)
    assert worst < 1e-13, worst
    rec.close()
    return nrec, nchk


@pytest.mark.parametrize("flavour,n,W,nshift,ri,P_", [
    (P.NS3D, 32, 8, (30, 2, 2), 0.0, 1.2),
    # ri = 1 here: with ri != 1 the reference's history norm (ri x^2 + y^2 for the
    # density) leaves a floor of (1 - ri) |Im rho|^2 under an exact recurrence
    (P.BOUS3D, 32, 8, (64, 1, 1), 1.0, 1.2),
    (P.NS2D, 128, 42, (30, 4, 1), 0.0, 1.0),
])
def test_planted_episode_matches_oracle(gpu, flavour, n, W, nshift, ri, P_):
    st = make(flavour, n, ri=ri)
    g = ogrid(flavour, n)
    kx, ky, kz, L = O.tables(g)
    nf = st.nf
    hist, _ = wave_history(kx, L, nf, 40, 0.1, P_, 7, nshift[0], t_break=3.05, seed=21 + nf)
    entry = lambda e: (st.set_state(hist[e]), hist[e])[1]
    nrec, nchk = compare_run(st, g, L, entry, 40, 10, 0.01,
                             dict(norm_width=W, nshift_x=nshift[0], nshift_y=nshift[1], nshift_z=nshift[2],
                                  store_size=100, threshold=1e-2, ri=ri), ri=ri)
    assert nrec == 1 and nchk >= 30
    st.close()


@pytest.mark.parametrize("flavour,n,nout,dt", [(P.NS3D, 32, 5, 0.01), (P.NS2D, 128, 10, 0.01),
                                               (P.BOUS3D, 32, 10, 0.01)])
def test_dns_history_matches_oracle(gpu, flavour, n, nout, dt):
    # a DNS stepped on the GPU; the oracle is fed the same states read back at each NOUT
    # block; a ring of 12 wraps during the 30 blocks (the Boussinesq rule compares
    # periods from 0.5 on: NOUT dt = 0.1 there); the threshold sits in the widest gap
    # of the middle half of the best residuals, so output sequences start and end
    st = make(flavour, n, dt=dt, v2=0.05 if flavour != P.NS2D else 1 / 40.0, ri=0.5)
    g = ogrid(flavour, n)
    kx, ky, kz, L = O.tables(g)
    if flavour == P.NS2D:
        s0 = P.ic_ns2d_cfg(st.grid, 5)[None]
    else:
        s0 = O.ic_solenoidal(g, 9, 0.5, rho_scale=0.1 if flavour == P.BOUS3D else None)
    st.set_state(s0)
    st.run(200)

    def entry(e):
        if e:
            st.run(nout)
        return st.get_state()

    opts = dict(store_size=12, threshold=10.0)
    if flavour == P.BOUS3D:
        opts["ri"] = 0.5
    # first pass: best residuals of each check, to place the threshold between two of them
    probe = P.Stepper(cfg=st.cfg)
    probe.set_state(st.get_state())
    rec = P.Recurrence(probe, nout=nout, **opts)
    best = []
    for e in range(1, 30):
        probe.run(nout)
        _, _, c = rec.update(e * nout, dt)
        if c["checked"] and c["compared"]:
            best.append(c["residual"])
    rec.close()
    probe.close()
    b = np.sort(best)
    gaps = np.diff(b) / b[1:]
    i = int(np.argmax(gaps[len(b) // 4:3 * len(b) // 4])) + len(b) // 4
    opts["threshold"] = 0.5 * (b[i] + b[i + 1])
    nrec, nchk = compare_run(st, g, L, entry, 30, nout, dt, opts, ri=opts.get("ri", 0.0))
    assert nchk >= 20
    st.close()


def test_config_size_256_check_against_direct_formula(gpu):
    # 256^3 with the reference's subset (W 8) and shifts (30x2x2): the residual table
    # of a check equals the direct norm of the difference on the subset (numpy)
    st = make(P.NS3D, 256, dt=0.002, v2=0.01)
    g = O.grid3(256, 256, 256)
    kx, ky, kz, L = O.tables(g)
    rng = np.random.default_rng(1)
    states = [(rng.standard_normal((3, g.nkt)) + 1j * rng.standard_normal((3, g.nkt))) * L for _ in range(4)]
    st.set_state(states[0])
    rec = P.Recurrence(st, nout=10, store_size=500, keep_states=0)
    for e in range(1, 4):
        st.set_state(states[e])
        rec.update(10 * e, 0.002)
    res, slots = rec.last()
    assert list(slots) == [1]
    W = 8
    idx = np.array([((g.ktz - W + iz) * g.ikty + g.ktz - W + iy) * g.iktx + ix
                    for iz in range(2 * W) for iy in range(2 * W) for ix in range(W)])
    zc, zj = states[2][:, idx], states[1][:, idx]
    SX = np.tile(2 * np.pi * np.arange(30) / 30, 4)
    SY = np.repeat(np.tile(2 * np.pi * np.arange(2) / 2, 2), 30)
    SZ = np.repeat(2 * np.pi * np.arange(2) / 2, 60)
    th = np.outer(SX, kx[idx]) + np.outer(SY, ky[idx]) + np.outer(SZ, kz[idx])
    D = np.exp(-1j * th)[:, None, :] * zc[None]
    direct = np.sum(np.abs(D - zj[None]) ** 2, axis=(1, 2)) / np.sum(np.abs(zc) ** 2 * L[idx][None])
    assert np.max(np.abs(res[0] ** 2 - direct)) < 1e-12
    rec.close()
    st.close()


def test_rejects_bad_options(gpu):
    st = make(P.NS3D, 32)
    with pytest.raises(P.PsgError, match="norm_width"):
        P.Recurrence(st, norm_width=11)
    with pytest.raises(P.PsgError, match="rule"):
        P.Recurrence(st, rule=2)
    rec = P.Recurrence(st, keep_states=0, store_size=4)
    with pytest.raises(P.PsgError, match="keep_states"):
        rec.state(0)
    rec.close()
    st.close()
