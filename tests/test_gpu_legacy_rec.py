"""Recurrence search through the legacy DNS symbols (RecFLAG = 1): the
3D Navier-Stokes and Boussinesq timestep_cuda_ calls write JOB0/UPO_ZK.out
and JOB0/UPOinfo_ts.dat; the expected records come from the oracle stepper
and the oracle search replaying the reference's NOUT blocks
(3D_NavierStokes/src/cuda.cu:114-139; 3D_Boussinesq/src/out_low.h:5-6 with
out_high.h:8-33).  The threshold is placed in the widest gap between the
checks' best residuals (from an oracle pre-pass), so that output sequences
start and end and rounding cannot move a decision."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_legacy import arr, bous_args, load, r, tables3

pytestmark = pytest.mark.gpu


def gap_threshold(best):
    """midpoint of the widest relative gap between the checks' best residuals
    (10000 marks a check without a candidate past the turning point)"""
    b = np.sort([x for x in best if x < 1000.0])
    assert len(b) >= 4
    gaps = np.diff(b) / b[1:]
    i = int(np.argmax(gaps))
    return 0.5 * (b[i] + b[i + 1]), gaps[i]


def read_upozk(path, nf, nkt):
    raw = np.fromfile(path, np.float64)
    rl = 7 + 2 * nf * nkt
    assert raw.size % rl == 0
    out = []
    for k in range(raw.size // rl):
        rec = raw[k * rl:(k + 1) * rl]
        out.append((rec[:7], rec[7:].view(np.complex128).reshape(nf, nkt)))
    return out


def check_files(recs, nf, nkt):
    got = read_upozk("JOB0/UPO_ZK.out", nf, nkt)
    assert len(got) == len(recs) >= 1
    for (h, s), (g, gs) in zip(got, recs):
        want = [g["time"], g["v2"], g["period"], g["sx"], g["sy"], g["sz"]]
        assert np.allclose(h[:6], want, rtol=1e-12, atol=1e-12)
        assert abs(h[6] - g["residual"]) < 1e-8 * g["residual"]
        assert np.abs(s - gs).max() < 1e-10 * np.abs(gs).max()
    lines = open("JOB0/UPOinfo_ts.dat").read().splitlines()
    assert lines[0].startswith("Guess No : Start time")
    assert len(lines) == len(recs) + 1
    for line, (g, _) in zip(lines[1:], recs):
        v = [float(x) for x in line.split()]
        assert int(v[0]) == g["index"]
        assert np.allclose(v[1:], [g["time"], g["period"], g["sx"], g["sy"], g["sz"], g["residual"]], rtol=1e-5)


def test_ns3d_dns_recflag_writes_oracle_guesses(gpu, tmp_path, monkeypatch):
    n = 32
    monkeypatch.chdir(tmp_path)
    os.makedirs("JOB0")
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    dt, v2, kf, nstop, nout, t0 = 0.005, 0.01, 4.0, 600, 10, 0.25
    nu, nuzn, nu1 = O.nutab(g, v2, dt)
    st0 = O.ic_ns3d_ref(g, 4242, 1.0)
    orc = O.NS3D(n, v2=v2, dt=dt, force=1, kf=kf)

    def replay(thr):
        rec = O.Rec(0, g, L, st0, store_size=500, norm_width=8, nshift=(30, 2, 2), nout=nout, dt=dt, tstart=t0,
                    threshold=thr, v2=v2)
        s, recs, best = orc.run(st0, nout + 1), [], []
        for NT in range(nout, nstop, nout):
            if NT > nout:
                s = orc.run(s, nout)
            gg, gs, c = rec.update(NT, dt, s, want_state=True)
            if c["checked"] and c["compared"]:
                best.append(c["residual"])
            if gg is not None:
                recs.append((gg, gs))
        return recs, best

    thr, gap = gap_threshold(replay(100.0)[1])
    assert gap > 1e-4
    recs, _ = replay(thr)
    x, e, z = (st0[i].copy() for i in range(3))
    nr = n ** 3
    xr, er, zr = np.zeros(nr), np.zeros(nr), np.zeros(nr)
    d = lambda v: C.c_double(v)
    i = lambda v: C.c_int(v)
    T, T0, amp, DT, res, KF, V2, al = d(0), d(t0), d(-0.25 * kf * kf), d(dt), d(thr), d(kf), d(v2), d(1.0)
    IKTX, IKTY, IKTZ, KTZ, NKT = i(g.iktx), i(g.ikty), i(g.iktz), i(g.ktz), i(g.nkt)
    NX = NY = NZ = i(n)
    NSTOP, NOUT, NOUTV, SF, RF, RANK = i(nstop), i(nout), i(100000), i(1), i(1), i(0)
    load("ns3d_dns").timestep_cuda_(
        arr(x), arr(e), arr(z), arr(xr), arr(er), arr(zr), arr(nuzn), arr(nu1), arr(kx), arr(ky), arr(kz), r(T),
        r(T0), r(amp), r(DT), r(res), r(KF), r(V2), r(al), arr(ikf), arr(ikn), arr(L), r(IKTX), r(IKTY), r(IKTZ),
        r(KTZ), r(NKT), r(NX), r(NY), r(NZ), r(NSTOP), r(NOUT), r(NOUTV), r(SF), r(RF), r(RANK))
    check_files(recs, 3, g.nkt)
    # stats.dat times: TSTART + 11 dt at the first block, then the check's reset
    # TSTART + (NT - NOUT) dt (functions.h:280 runs before computeStats)
    # (the last line is finalise.h's closing computeStats)
    t = [float(l.split()[0]) for l in open("JOB0/stats.dat").read().splitlines()]
    assert abs(t[0] - (t0 + (nout + 1) * dt)) < 1e-6
    assert np.allclose(t[1:-1], [t0 + (NT - nout) * dt for NT in range(2 * nout, nstop, nout)], rtol=1e-6)


def test_bous_dns_recflag_writes_oracle_guesses(gpu, tmp_path, monkeypatch):
    n = 32
    monkeypatch.chdir(tmp_path)
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    v2, sc, ri, theta, dt, kf, Tstep, nout = 1 / 100.0, 1.0, 1.0, np.pi / 2, 0.01, 4.0, 6.0, 5
    st0 = O.ic_solenoidal(g, 63, 0.5, 0.1)
    orc = O.BOUS(n, v2=v2, sc=sc, ri=ri, theta=theta, dt=dt)
    small, tout = 1e-10, nout * dt

    def crossing(tme):
        crit = tout * int((tme + small) / tout)
        return (tme - dt - crit + small) * (tme + small - crit) < 0.0

    def replay(thr):
        # the reference's loop: tcheck at the top of a step, the block after it
        rec = O.Rec(1, g, L, st0, store_size=500, norm_width=8, nshift=(64, 1, 1), nout=nout, dt=dt, tstart=0.0,
                    threshold=thr, ri=ri, v2=v2)
        s, tme, k, pend, recs, best = st0, 0.0, 0, 0, [], []
        while tme <= Tstep:
            hi = crossing(tme)
            pend += 1
            tme += dt
            k += 1
            if hi:
                s = orc.run(s, pend)
                pend = 0
                gg, gs, c = rec.update(k, dt, s, want_state=True)
                if c["checked"] and c["compared"]:
                    best.append(c["residual"])
                if gg is not None:
                    recs.append((gg, gs))
        return recs, best

    thr, gap = gap_threshold(replay(100.0)[1])
    assert gap > 1e-4
    recs, _ = replay(thr)
    st = [st0[i].copy() for i in range(4)]
    T, T0 = C.c_double(0.0), C.c_double(0.0)
    a, keep = bous_args(g, n, st, T, T0, Tstep, -kf / 2, dt, kf, v2, ri, sc, theta, L, kx, ky, kz, ikf, ikn,
                        nout=nout, noutv=1000, stats=0, rec=1, thr=thr)
    load("bous_dns").timestep_cuda_(*a)
    check_files(recs, 4, g.nkt)
