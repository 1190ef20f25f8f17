"""3D Navier-Stokes DNS end-of-run outputs through the legacy symbol
(statsFLAG = 1): the velocity variances accumulated every NOUT steps
(avgVelocity, 3D_NavierStokes/src/cuda.cu:70), the last computeStats line,
avgs.dat, means.dat and pdfs.dat, against a replay of the reference's loop
on the oracle (stepper, computeStats sums, physical fields) with the
reference's scalar bookkeeping restated here (functions.h:128-209,
finalise.h:16-167; its file-scope min/max start at zero)."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_legacy import arr, load, r, tables3

pytestmark = pytest.mark.gpu


def replay(n, st0, v2, dt, kf, nstop, nout, amp):
    g = O.grid3(n, n, n)
    orc = O.NS3D(n, v2=v2, dt=dt, force=1, kf=kf)
    t = np.zeros((3, g.nkt), np.complex128)
    rms = np.zeros((3, n, n, n))
    nstats = nstop // nout + 1
    arrs = {k: np.zeros(nstats) for k in ("diss", "energy", "input")}
    b = dict(dissMax=0.0, dissMin=0.0, enerMax=0.0, enerMin=0.0, inputMax=0.0, inputMin=0.0)
    tc, tme, s = 0, 0.0, st0.copy()
    lines = []

    def compute(NT):
        nonlocal tc
        raw = O.stats_ns3d(g, s, t, tc, amp, kf)
        inp = raw[4] * v2 * kf * kf
        diss = raw[0] * v2 * v2 * 8.0 * kf * kf
        energy = raw[1]
        for k, v in (("diss", diss), ("ener", energy), ("input", inp)):
            b[k + "Max"] = max(v, b[k + "Max"])
            b[k + "Min"] = min(v, b[k + "Min"])
        istat = NT // nout - 1
        arrs["diss"][istat], arrs["energy"][istat], arrs["input"][istat] = diss, energy, inp
        lines.append([tme, energy, diss, inp])
        tc += 1

    for NT in range(nstop):
        if NT % nout == 0:
            u = orc.physical(s)[3:6]
            rms = (rms * tc + u ** 2) / (tc + 1.0)
        s = orc.run(s, 1)
        tme += dt
        if NT % nout == 0 and NT != 0:
            compute(NT)
    compute(nstop + 1)
    return dict(state=s, tavg=t, rms=rms, arrs=arrs, b=b, lines=np.array(lines), nstats=nstats, orc=orc)


def pdfs(want):
    b, a, ns = want["b"], want["arrs"], want["nstats"]
    nbin, corr = 100, 1.0000001
    d = {k: (b[k + "Max"] * corr - b[k + "Min"]) / nbin for k in ("diss", "ener", "input")}
    hist = {k: np.zeros(nbin + 2) for k in d}
    for i in range(ns - 1):
        for k, src in (("ener", "energy"), ("diss", "diss"), ("input", "input")):
            j = int((a[src][i] - b[k + "Min"]) / d[k]) + 1
            hist[k][j] += 1
    rows = []
    for ib in range(nbin):
        row = []
        for k in ("ener", "diss", "input"):
            row += [b[k + "Min"] + (ib - 0.5) * d[k], hist[k][ib] / ((ns - 1) * d[k])]
        rows.append(row)
    return np.array(rows)


def test_ns3d_dns_end_of_run_files_match_oracle_replay(gpu, tmp_path, monkeypatch):
    n, dt, v2, kf, nstop, nout = 32, 0.005, 0.01, 4.0, 20, 5
    amp = -0.25 * kf * kf
    monkeypatch.chdir(tmp_path)
    os.makedirs("JOB0")
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    nu, nuzn, nu1 = O.nutab(g, v2, dt)
    st0 = O.ic_ns3d_ref(g, 4243, 1.0)
    want = replay(n, st0, v2, dt, kf, nstop, nout, amp)
    x, e, z = (st0[i].copy() for i in range(3))
    nr = n ** 3
    xr, er, zr = np.zeros(nr), np.zeros(nr), np.zeros(nr)
    d = lambda v: C.c_double(v)
    i = lambda v: C.c_int(v)
    T, T0, AMP, DT, res, KF, V2, al = d(0), d(0), d(amp), d(dt), d(1.0), d(kf), d(v2), d(1.0)
    IKTX, IKTY, IKTZ, KTZ, NKT = i(g.iktx), i(g.ikty), i(g.iktz), i(g.ktz), i(g.nkt)
    NX = NY = NZ = i(n)
    NSTOP, NOUT, NOUTV, SF, RF, RANK = i(nstop), i(nout), i(100000), i(1), i(0), i(0)
    load("ns3d_dns").timestep_cuda_(
        arr(x), arr(e), arr(z), arr(xr), arr(er), arr(zr), arr(nuzn), arr(nu1), arr(kx), arr(ky), arr(kz), r(T),
        r(T0), r(AMP), r(DT), r(res), r(KF), r(V2), r(al), arr(ikf), arr(ikn), arr(L), r(IKTX), r(IKTY), r(IKTZ),
        r(KTZ), r(NKT), r(NX), r(NY), r(NZ), r(NSTOP), r(NOUT), r(NOUTV), r(SF), r(RF), r(RANK))
    got = np.stack([x, e, z])
    assert np.abs(got - want["state"]).max() < 1e-11 * np.abs(want["state"]).max()
    # stats.dat: NT = 5, 10, 15 in the loop, then the last computeStats of finalise.h
    st = np.array([[float(v) for v in l.split()] for l in open("JOB0/stats.dat").read().splitlines()])
    assert st.shape == (4, 6)
    assert np.allclose(st[:, [0, 1, 3, 5]], want["lines"], rtol=2e-6, atol=1e-300)
    # avgs.dat: five time-stamped fields of the time-averaged state and the variances
    raw = np.fromfile("JOB0/avgs.dat", np.float64).reshape(5, nr + 1)
    assert np.all(raw[:, 0] == nstop * dt)
    pt = want["orc"].physical(want["tavg"])
    rms = want["rms"]
    fields = [pt[3], pt[4], pt[5], rms[1], rms[0] - pt[3] ** 2]
    for k, f in enumerate(fields):
        assert np.abs(raw[k, 1:] - f.ravel()).max() <= 1e-10 * np.abs(f).max(), k
    # means.dat: (y, z, <u>, <v>, x-mean of <u^2>-<u>^2, x-mean of <v^2>) per (z, y)
    me = np.array([[float(v) for v in l.split()] for l in open("JOB0/means.dat").read().splitlines()])
    assert me.shape == (n * n, 6)
    cols = [fields[0], fields[1], fields[4], fields[3]]
    for q, f in enumerate(cols):
        assert np.allclose(me[:, 2 + q], f.mean(axis=2).ravel(), rtol=1e-10, atol=1e-13), q
    assert np.allclose(me[:n, 0], np.arange(n) * 2 * np.pi / n) and np.all(me[:n, 1] == 0.0)
    # pdfs.dat: 100 bins of energy, dissipation and input over the stored computeStats values
    assert want["b"]["dissMax"] - want["b"]["dissMin"] > 0.001
    pd = np.array([[float(v) for v in l.split()] for l in open("JOB0/pdfs.dat").read().splitlines()])
    assert pd.shape == (100, 6)
    assert np.allclose(pd, pdfs(want), rtol=1e-6, atol=1e-12)
