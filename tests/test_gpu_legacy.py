"""The reference's own call sequence through the legacy drop-in symbols:
timestep_cuda_ called Fortran-style (every argument by address) with the
caller-built tables, compared against the FP64 oracle."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
LIBDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "psgpu_amd", "_lib")


def load(name):
    return C.CDLL(os.path.join(LIBDIR, "libpsg_legacy_%s.so" % name))


def r(x):
    return C.byref(x)


def arr(a):
    return a.ctypes.data_as(C.c_void_p)


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def tables3(n):
    g = O.grid3(n, n, n)
    kx, ky, kz, L = O.tables(g)
    ikf, ikn = O.ikf_ikn(g)
    return g, kx, ky, kz, L.astype(np.int32), ikf, ikn


def test_legacy_ns3d_gmres_matches_oracle(gpu, monkeypatch):
    n = 32
    monkeypatch.setenv("PSG_NX", str(n)); monkeypatch.setenv("PSG_NY", str(n)); monkeypatch.setenv("PSG_NZ", str(n))
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    dt, v2, kf = 0.002, 1 / 40.0, 4
    nu, nuzn, nu1 = O.nutab(g, v2, dt)
    st = O.ic_ns3d_ref(g, 837920, 1.0)
    x, e, z = (st[i].copy() for i in range(3))   # copies: the call writes in place
    T, T0, amp, DT, V2 = C.c_double(0), C.c_double(1.5), C.c_double(-0.25 * kf), C.c_double(dt), C.c_double(v2)
    KF, NST = C.c_int(kf), C.c_int(30)
    lib = load("ns3d_gmres")
    lib.timestep_cuda_(arr(x), arr(e), arr(z), arr(nuzn), arr(nu1), arr(kx), arr(ky), arr(kz), r(T), r(T0), r(amp),
                       r(DT), r(KF), r(V2), arr(ikf), arr(ikn), arr(L), r(NST))
    want = O.NS3D(n, v2=v2, dt=dt, force=2, kf=kf).run(st, 30)
    assert rel(np.stack([x, e, z]), want) < 1e-11
    assert abs(T.value - (1.5 + 30 * dt)) < 1e-14
    # a second call reuses the cached context; a shortened last step changes dt
    DT2 = C.c_double(0.0007)
    NST1 = C.c_int(1)
    nuzn2, nu12 = O.nutab(g, v2, 0.0007)[1:]
    lib.timestep_cuda_(arr(x), arr(e), arr(z), arr(nuzn2), arr(nu12), arr(kx), arr(ky), arr(kz), r(T), r(T0),
                       r(amp), r(DT2), r(KF), r(V2), arr(ikf), arr(ikn), arr(L), r(NST1))
    want2 = O.NS3D(n, v2=v2, dt=0.0007, force=2, kf=kf).run(want, 1)
    assert rel(np.stack([x, e, z]), want2) < 1e-11


def test_legacy_ns3d_dns_outputs(gpu, tmp_path, monkeypatch):
    n = 32
    monkeypatch.chdir(tmp_path)
    os.makedirs("JOB0")
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    dt, v2, kf = 0.005, 0.01, 4.0
    nu, nuzn, nu1 = O.nutab(g, v2, dt)
    st = O.ic_ns3d_ref(g, 4242, 1.0)
    x, e, z = (st[i].copy() for i in range(3))   # copies: the call writes in place
    nr = n ** 3
    xr, er, zr = np.zeros(nr), np.zeros(nr), np.zeros(nr)
    d = lambda v: C.c_double(v)
    i = lambda v: C.c_int(v)
    T, T0, amp, DT, res, KF, V2, al = d(0), d(0), d(-0.25 * kf * kf), d(dt), d(1.0), d(kf), d(v2), d(1.0)
    IKTX, IKTY, IKTZ, KTZ, NKT = i(g.iktx), i(g.ikty), i(g.iktz), i(g.ktz), i(g.nkt)
    NX = NY = NZ = i(n)
    NSTOP, NOUT, NOUTV, SF, RF, RANK = i(20), i(5), i(10), i(1), i(0), i(0)
    load("ns3d_dns").timestep_cuda_(
        arr(x), arr(e), arr(z), arr(xr), arr(er), arr(zr), arr(nuzn), arr(nu1), arr(kx), arr(ky), arr(kz), r(T),
        r(T0), r(amp), r(DT), r(res), r(KF), r(V2), r(al), arr(ikf), arr(ikn), arr(L), r(IKTX), r(IKTY), r(IKTZ),
        r(KTZ), r(NKT), r(NX), r(NY), r(NZ), r(NSTOP), r(NOUT), r(NOUTV), r(SF), r(RF), r(RANK))
    orc = O.NS3D(n, v2=v2, dt=dt, force=1, kf=kf)
    assert rel(np.stack([x, e, z]), orc.run(st, 20)) < 1e-11
    # raw dumps at NT=0 and NT=10: physical vorticity of the state before that step
    p0 = np.fromfile("JOB0/xii001.raw")
    p1 = np.fromfile("JOB0/zet002.raw")
    assert rel(p0, orc.physical(st)[0].ravel()) < 1e-12
    assert rel(p1, orc.physical(orc.run(st, 10))[2].ravel()) < 1e-12
    assert rel(zr, p1) == 0.0
    lines = open("JOB0/stats.dat").read().strip().splitlines()
    assert len(lines) == 4                    # after steps NT = 5, 10, 15, and finalise.h's last one
    t, en = map(float, lines[0].split()[:2])
    assert abs(t - 6 * dt) < 1e-12            # tme after iteration NT=5 (6 steps)
    assert abs(en - orc.diag(orc.run(st, 6))[0]) < 1e-6 * en


def test_legacy_ns2d_gmres_matches_oracle(gpu):
    n = 64
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    ikf, ikn = O.ikf_ikn(g)
    z0 = O.ic_ns2d_ref(g, 837920, 1.0)
    z = z0.copy()
    zr = np.zeros(n * n)
    d = lambda v: C.c_double(v)
    i = lambda v: C.c_int(v)
    T, T0, TN, DT, al, V2 = d(9.0), d(0.0), d(5.0), d(0.005), d(1.0), d(1 / 40.0)
    load("ns2d_gmres").timestep_cuda_(arr(z), arr(zr), arr(kx), arr(ky), r(T), r(T0), r(TN), r(DT), r(al), r(V2),
                                      arr(ikf), arr(ikn), r(i(g.iktx)), r(i(g.ikty)), r(i(g.kty)), r(i(n)),
                                      r(i(n)), r(i(40)), arr(L.astype(np.int32)))
    want = O.NS2D(n, v2=1 / 40.0, dt=0.005).run(z0, 40)
    assert rel(z, want) < 1e-11
    assert abs(T.value - 40 * 0.005) < 1e-13


def test_legacy_euler_matches_oracle(gpu, tmp_path, monkeypatch):
    n = 32
    monkeypatch.chdir(tmp_path)
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    st = O.ic_solenoidal(g, 77, 0.5)
    x, e, z = (st[i].copy() for i in range(3))   # copies: the call writes in place
    nr = n ** 3
    xr, er, zr = np.zeros(nr), np.zeros(nr), np.zeros(nr)
    d = lambda v: C.c_double(v)
    i = lambda v: C.c_int(v)
    T, T0, DT, al = d(0), d(0), d(0.005), d(1.0)
    load("euler").timestep_cuda_(arr(x), arr(e), arr(z), arr(xr), arr(er), arr(zr), arr(kx), arr(ky), arr(kz),
                                 r(T), r(T0), r(DT), r(al), arr(ikf), arr(ikn), arr(L), r(i(g.iktx)), r(i(g.ikty)),
                                 r(i(g.iktz)), r(i(g.ktz)), r(i(g.nkt)), r(i(n)), r(i(n)), r(i(n)), r(i(12)),
                                 r(i(4)), r(i(100)), r(i(0)))
    want = O.NS3D(n, dt=0.005, flavour=1, force=0).run(st, 12)
    assert rel(np.stack([x, e, z]), want) < 1e-11


def bous_args(g, n, st, T, T0, Tstep, amp, dt, kf, v2, ri, sc, theta, L, kx, ky, kz, ikf, ikn,
              nout=1, noutv=1, stats=0, rec=0, adapt=0, rank=0, thr=1.0):
    """The 37-argument Boussinesq list (3D_Boussinesq/src/cuda.cu:21-58), by address."""
    d = lambda v: C.c_double(v)
    i = lambda v: C.c_int(v)
    keep = [d(Tstep), d(amp), d(dt), d(thr), d(kf), d(v2), d(ri), d(sc), d(theta), d(1.0), d(0.0),
            i(g.iktx), i(g.ikty), i(g.iktz), i(g.ktz), i(g.nkt), i(n), i(n), i(n), i(nout), i(noutv), i(stats),
            i(rec), i(adapt), i(rank)]
    a = [arr(st[0]), arr(st[1]), arr(st[2]), arr(st[3]), arr(kx), arr(ky), arr(kz), r(T), r(T0)]
    a += [r(v) for v in keep[:11]] + [arr(ikf), arr(ikn), arr(L)] + [r(v) for v in keep[11:]]
    return a, keep


def test_legacy_bous_gmres_and_timederiv(gpu):
    n = 32
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    v2, sc, ri, theta, dt, kf, Tstep = 1 / 100.0, 1.0, 1.0, np.pi / 2, 0.004, 4.0, 0.1
    st0 = O.ic_solenoidal(g, 61, 0.5, 0.1)
    st = [st0[i].copy() for i in range(4)]
    T, T0 = C.c_double(0.0), C.c_double(0.0)
    lib = load("bous_gmres")
    a, keep = bous_args(g, n, st, T, T0, Tstep, -kf / 2, dt, kf, v2, ri, sc, theta, L, kx, ky, kz, ikf, ikn)
    lib.timestep_cuda_(*a)
    # the driver's loop (src_gmres/cuda.cu:188-262) replayed on the host
    tme, nst = 0.0, 0
    while Tstep - tme > dt:
        tme += dt
        nst += 1
    orc = O.BOUS(n, v2=v2, sc=sc, ri=ri, theta=theta, dt=dt)
    want = orc.run(st0, nst)
    if tme != Tstep:
        want = O.BOUS(n, v2=v2, sc=sc, ri=ri, theta=theta, dt=Tstep - tme).run(want, 1)
    assert rel(np.stack(st), want) < 1e-11
    assert abs(T.value - nst * dt) < 1e-14
    # timederiv_cuda_ overwrites the state with dz/dt
    st2 = [st0[i].copy() for i in range(4)]
    a, keep = bous_args(g, n, st2, T, T0, Tstep, -kf / 2, dt, kf, v2, ri, sc, theta, L, kx, ky, kz, ikf, ikn)
    lib.timederiv_cuda_(*a)
    assert rel(np.stack(st2), orc.rhs(st0)) < 1e-13


def test_legacy_bous_dns_steps_and_dumps(gpu, tmp_path, monkeypatch):
    n = 32
    monkeypatch.chdir(tmp_path)
    g, kx, ky, kz, L, ikf, ikn = tables3(n)
    v2, sc, ri, theta, dt, kf, Tstep = 1 / 100.0, 1.0, 1.0, np.pi / 2, 0.004, 4.0, 0.06
    st0 = O.ic_solenoidal(g, 62, 0.5, 0.1)
    st = [st0[i].copy() for i in range(4)]
    T, T0 = C.c_double(0.0), C.c_double(0.0)
    a, keep = bous_args(g, n, st, T, T0, Tstep, -kf / 2, dt, kf, v2, ri, sc, theta, L, kx, ky, kz, ikf, ikn,
                        nout=5, noutv=10, stats=1)
    load("bous_dns").timestep_cuda_(*a)
    # while (tme <= Tstep) with dt = Tstep/(Tstep/DELT) (init.h:24-29, cuda.cu:68-99)
    nstop = Tstep / dt
    d = Tstep / nstop if nstop * dt != Tstep else dt
    tme, nst = 0.0, 0
    while tme <= Tstep:
        tme += d
        nst += 1
    orc = O.BOUS(n, v2=v2, sc=sc, ri=ri, theta=theta, dt=d)
    assert rel(np.stack(st), orc.run(st0, nst)) < 1e-11
    assert abs(T.value - tme) < 1e-14
    # first dump: velocity and density of the initial state, x fastest
    ns = O.NS3D(n)
    phys = ns.physical(st0[:3])
    zero = np.zeros_like(st0[0])
    rho = ns.physical(np.stack([st0[3], zero, zero]))[0]
    for name, want in (("U", phys[3]), ("V", phys[4]), ("W", phys[5]), ("rho", rho)):
        got = np.fromfile("JOB0/%s001.raw" % name)
        assert rel(got, want.ravel()) < 1e-12, name
    assert os.path.exists("JOB0/U002.raw")


def test_bous_physical_fields(gpu):
    import psgpu_amd as P
    n = 32
    g = O.grid3(n, n, n)
    st = O.ic_solenoidal(g, 63, 0.5, 0.1)
    with P.Stepper(flavour="bous", n=n, v2=1 / 100.0, ri=1.0) as s:
        s.set_state(st)
        got = s.physical()
    ns = O.NS3D(n)
    zero = np.zeros_like(st[0])
    assert rel(got[:6], ns.physical(st[:3])) < 1e-13
    assert rel(got[6], ns.physical(np.stack([st[3], zero, zero]))[0]) < 1e-13
