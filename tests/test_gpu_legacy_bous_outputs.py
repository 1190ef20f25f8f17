"""Boussinesq production outputs through the legacy DNS symbol (statsFLAG = 1):
stats.dat, engy.dat, RiG.dat, spec.dat, avgs.dat, means.dat and sweep.dat
against a replay of the reference's loop on the oracle (stepper, computeStats
sums, RiG plane sums, spectra, physical fields) with the reference's scalar
bookkeeping (3D_Boussinesq/src/functions.h:361-410, out_low.h:8-76,
finalise.h:17-200) restated here; and the Dtarget throttle against the same
replay with the amplitude changed between steps."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_legacy import bous_args, load, tables3

pytestmark = pytest.mark.gpu

SMALL = 1e-10


def crossing(tme, dt, period):
    crit = period * int((tme + SMALL) / period)
    return (tme - dt - crit + SMALL) * (tme + SMALL - crit) < 0.0


def replay(n, st0, v2, ri, theta, dt, kf, Tstep, nout, amp0, dtarget=0.0):
    g = O.grid3(n, n, n)
    ns = O.NS3D(n)
    zero = np.zeros(g.nkt, np.complex128)
    nstop = Tstep / dt                      # init.h:24-29
    dt = Tstep / nstop if nstop * dt != Tstep else dt
    tout = nout * dt
    tme, s, amp = 0.0, st0.copy(), amp0
    t = np.zeros((4, g.nkt), np.complex128)
    rms = np.zeros((3, n, n, n))
    b = dict(dissMax=0.0, dissMin=1e4, inputMax=0.0, inputMin=1e4, enerMax=0.0, enerMin=1e4, enerAvg=0.0,
             dissAvg=0.0, peAvg=0.0, pedAvg=0.0, mixAvg=0.0, inputAvg=0.0, eturbAvg=0.0, dturbAvg=0.0,
             bfluxAvg=0.0, RiGMax=0.0)
    tc = 0
    stats, engy, rig = [], [], []

    def compute(s, tme):
        nonlocal tc, amp
        raw = O.stats_bous(g, s, t, tc, amp, kf, ri, theta)
        diss, energy, eturb, dturb, inp, pe, ped, ped1, ped2, peturb, pedturb, bflux = raw[[0, 1, 2, 3, 4, 5, 6, 7, 8,
                                                                                             9, 10, 11]]
        if dtarget != 0.0:
            amp = amp * (dtarget - bflux) / inp
        b["dissAvg"] = (b["dissAvg"] * tc + diss) / (tc + 1.0)
        b["enerAvg"] = (b["enerAvg"] * tc + energy) / (tc + 1.0)
        b["peAvg"] = (b["peAvg"] * tc + pe) / (tc + 1.0)
        b["pedAvg"] = (b["pedAvg"] * tc + ped) / (tc + 1.0)
        b["inputAvg"] = (b["inputAvg"] * tc + inp) / (tc + 1.0)
        b["bfluxAvg"] = (b["bfluxAvg"] * tc + bflux) / (tc + 1.0)
        b["mixAvg"] = ((ped / (diss / ri + ped)) * tc + b["mixAvg"]) / (tc + 1.0)
        inp *= v2 * kf * kf * 2.0
        diss *= v2 * v2 * 4.0 * kf * kf
        dturb = (b["dissAvg"] - dturb) * v2 * v2 * 4.0 * kf * kf
        eturb = b["enerAvg"] - eturb
        peturb = b["peAvg"] - peturb
        pedturb = b["pedAvg"] - pedturb
        tc += 1
        stats.append([tme, energy, eturb, diss, dturb, inp, 0.0])
        engy.append([tme, bflux, pe, peturb, ped, pedturb, ped1, ped2])

    while tme <= Tstep:
        tcheck = crossing(tme, dt, tout)
        if tcheck:
            ph = ns.physical(s[:3])
            ro = ns.physical(np.stack([s[3], zero, zero]))[0]
            rms = (rms * tc + np.stack([ph[3], ph[4], ph[5]]) ** 2) / (tc + 1.0)
            na, sa = O.rig(n, n, n, theta, ph[3], ph[4], ph[5], ro)
            b["RiGMax"] = max([0.0] + list(na / n / sa / n))
            rig.append([(tme, i, na[i] / (n * n), sa[i] / (n * n)) for i in range(len(na))])
        s = O.BOUS(n, v2=v2, ri=ri, theta=theta, dt=dt, kf=kf, ampfor=amp).run(s, 1)
        tme += dt
        if tcheck:
            compute(s, tme)
    compute(s, tme)   # finalise.h:20-22
    return dict(state=s, stats=np.array(stats), engy=np.array(engy), rig=rig, tavg=t, rms=rms, b=b, g=g, amp=amp)


def run_shim(n, st0, v2, ri, theta, dt, kf, Tstep, nout, amp0, dtarget=0.0):
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    st = [st0[i].copy() for i in range(4)]
    T, T0 = C.c_double(0.0), C.c_double(0.0)
    a, keep = bous_args(g, n, st, T, T0, Tstep, amp0, dt, kf, v2, ri, 1.0, theta, L, kx, ky, kz, ikf, ikn,
                        nout=nout, noutv=1000, stats=1)
    keep[10].value = dtarget
    load("bous_dns").timestep_cuda_(*a)
    return np.stack(st), T.value


def read_rows(path):
    rows = []
    for line in open(path).read().splitlines():
        if line.strip():
            rows.append([float(x) for x in line.split()])
    return np.array(rows)


@pytest.mark.parametrize("theta", [np.pi / 2, np.pi / 8])
def test_bous_stats_files_match_oracle_replay(gpu, tmp_path, monkeypatch, theta):
    n, v2, ri, dt, kf, Tstep, nout = 32, 0.01, 0.7, 0.004, 4.0, 0.2, 5
    monkeypatch.chdir(tmp_path)
    g = O.grid3(n, n, n)
    st0 = O.ic_solenoidal(g, 64, 0.5, 0.1)
    want = replay(n, st0, v2, ri, theta, dt, kf, Tstep, nout, -kf / 2)
    got, Tend = run_shim(n, st0, v2, ri, theta, dt, kf, Tstep, nout, -kf / 2)
    assert np.abs(got - want["state"]).max() < 1e-11 * np.abs(want["state"]).max()
    st = read_rows("JOB0/stats.dat")
    assert st.shape == want["stats"].shape
    assert np.allclose(st, want["stats"], rtol=2e-6, atol=1e-300)
    en = read_rows("JOB0/engy.dat")
    assert np.allclose(en, want["engy"], rtol=2e-6, atol=1e-12)
    rig = read_rows("JOB0/RiG.dat")
    flat = np.array([r for blk in want["rig"] for r in blk])
    assert rig.shape == flat.shape and np.allclose(rig, flat, rtol=2e-6, atol=1e-12)
    # spectra of the time-averaged vorticity (finalise.h:24-60)
    h, v, size = O.spectra(g, want["tavg"], theta)
    sp = read_rows("JOB0/spec.dat")
    assert sp.shape == (size, 3)
    assert np.allclose(sp[:, 1], h[:size], rtol=2e-6, atol=1e-300) and np.allclose(sp[:, 2], v[:size], rtol=2e-6,
                                                                                    atol=1e-300)
    # avgs.dat: seven time-stamped fields; means.dat: x-means per (y, z)
    nr = n ** 3
    raw = np.fromfile("JOB0/avgs.dat", np.float64).reshape(7, nr + 1)
    assert np.all(raw[:, 0] == Tend)
    ns = O.NS3D(n)
    zero = np.zeros(g.nkt, np.complex128)
    pt = ns.physical(want["tavg"][:3])
    rt = ns.physical(np.stack([want["tavg"][3], zero, zero]))[0]
    fields = [rt, pt[3], pt[4], pt[5], want["rms"][1], want["rms"][2], want["rms"][0] - pt[3] ** 2]
    for k, f in enumerate(fields):
        assert np.abs(raw[k, 1:] - f.ravel()).max() <= 1e-10 * max(1e-300, np.abs(f).max()), k
    me = read_rows("JOB0/means.dat")
    assert me.shape == (n * n, 6)
    assert np.allclose(me[:, 2], pt[3].mean(axis=2).ravel(), rtol=1e-12, atol=1e-14)
    b = want["b"]
    sw = read_rows("sweep.dat")[-1]
    assert np.allclose(sw, [1 / v2, ri, 1.0, b["dissAvg"], b["inputAvg"], b["mixAvg"], b["RiGMax"], b["pedAvg"]],
                       rtol=1e-11)
    assert os.path.exists("JOB0/pdfs.dat")


def test_bous_dtarget_throttle_matches_oracle_replay(gpu, tmp_path, monkeypatch):
    n, v2, ri, theta, dt, kf, Tstep, nout = 32, 0.01, 1.0, np.pi / 2, 0.004, 4.0, 0.12, 5
    monkeypatch.chdir(tmp_path)
    g = O.grid3(n, n, n)
    st0 = O.ic_solenoidal(g, 65, 0.5, 0.1)
    want = replay(n, st0, v2, ri, theta, dt, kf, Tstep, nout, -kf / 2, dtarget=0.05)
    assert abs(want["amp"] + kf / 2) > 1e-6          # the throttle moved the amplitude
    got, _ = run_shim(n, st0, v2, ri, theta, dt, kf, Tstep, nout, -kf / 2, dtarget=0.05)
    assert np.abs(got - want["state"]).max() < 1e-10 * np.abs(want["state"]).max()
