"""GPU parity: the HIP path (through the C-ABI) against the FP64 CPU oracle on
the same seeded inputs.  Tolerances: plain FP64 form <= 1e-12 relative for
energy, enstrophy and shell spectra over 100 steps (the benchmark's stated
bar), <= 1e-13 for single right-hand-side evaluations; FP32-bulk is checked
against its own stated bound."""
import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def diag_rel(dg, do):
    e = abs(dg[0] - do[0]) / abs(do[0])
    z = abs(dg[1] - do[1]) / abs(do[1])
    sp = rel(dg[2], do[2])
    return e, z, sp


def test_nonlin_3d(gpu):
    n = 32
    st = O.ic_solenoidal(O.grid3(n, n, n), 3)
    with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, force="kolmogorov") as s:
        got = s.nonlin(st)
    want = O.NS3D(n, v2=0.01, dt=0.002, force=3).nonlin(st)
    assert rel(got, want) < 1e-13


def test_nonlin_3d_aspect(gpu):
    # NX = NZ/alpha box as in 3D_Boussinesq (alpha = 0.5)
    g = O.grid3(64, 32, 32, 0.5)
    st = O.ic_solenoidal(g, 4)
    with P.Stepper(flavour="ns3d", nx=64, ny=32, nz=32, alpha=0.5, force="dns") as s:
        got = s.nonlin(st)
    want = O.NS3D(32, alpha=0.5, force=1, nx=64, ny=32, nz=32).nonlin(st)
    assert rel(got, want) < 1e-13


def test_nonlin_2d(gpu):
    g = O.grid2(64, 64)
    z = O.ic_ns2d_cfg(g, 5)
    with P.Stepper(flavour="ns2d", n=64, v2=1 / 40, dt=0.005, force="kolmogorov", kf=4.0, ampfor=-2.0) as s:
        got = s.nonlin(z[None])[0]
    want = O.NS2D(64).nonlin(z)
    assert rel(got, want) < 1e-13


def test_physical_fields_3d(gpu):
    n = 32
    st = O.ic_solenoidal(O.grid3(n, n, n), 8)
    with P.Stepper(flavour="ns3d", n=n) as s:
        s.set_state(st)
        got = s.physical()
    want = O.NS3D(n).physical(st)
    assert rel(got, want) < 1e-13


@pytest.mark.parametrize("force", ["kolmogorov", "dns", "gmres"])
def test_ns3d_parity_100_steps(gpu, force):
    n = 32
    g = O.grid3(n, n, n)
    st = O.ic_solenoidal(g, 21) if force == "kolmogorov" else O.ic_ns3d_ref(g, 837920, 1.0)
    fmode = {"kolmogorov": 3, "dns": 1, "gmres": 2}[force]
    orc = O.NS3D(n, v2=0.01, dt=0.002, force=fmode)
    want = orc.run(st, 100)
    with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, force=force) as s:
        s.set_state(st)
        s.run(100)
        got = s.get_state()
        dg = s.diag()
    e, z, sp = diag_rel(dg, orc.diag(want))
    assert e < 1e-12 and z < 1e-12 and sp < 1e-12, (e, z, sp)
    assert rel(got, want) < 1e-11


def test_ns3d_parity_64(gpu):
    n = 64
    g = O.grid3(n, n, n)
    st = O.ic_solenoidal(g, 22)
    orc = O.NS3D(n, v2=0.01, dt=0.002, force=3)
    with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002) as s:
        s.set_state(st)
        cur = st
        for chunk in range(4):
            s.run(25)
            cur = orc.run(cur, 25)
            e, z, sp = diag_rel(s.diag(), orc.diag(cur))
            assert e < 1e-12 and z < 1e-12 and sp < 1e-12, (chunk, e, z, sp)


def test_euler_parity(gpu):
    n = 32
    st = O.ic_solenoidal(O.grid3(n, n, n), 31)
    orc = O.NS3D(n, dt=0.005, flavour=1, force=0)
    want = orc.run(st, 50)
    with P.Stepper(flavour="euler", n=n, dt=0.005, force="none") as s:
        s.set_state(st)
        s.run(50)
        got = s.get_state()
    assert rel(got, want) < 1e-11
    e, z, sp = diag_rel((*orc.diag(got)[:2], orc.diag(got)[2]), orc.diag(want))
    assert e < 1e-12 and z < 1e-12


def test_ns2d_parity_256(gpu):
    n = 256
    g = O.grid2(n, n)
    z0 = O.ic_ns2d_cfg(g, 837920, 1.0)
    orc = O.NS2D(n, v2=1 / 40, dt=0.005)
    with P.Stepper(flavour="ns2d", n=n, v2=1 / 40, dt=0.005, force="kolmogorov", kf=4.0, ampfor=-2.0) as s:
        s.set_state(z0[None])
        cur = z0
        for chunk in range(4):
            s.run(25)
            cur = orc.run(cur, 25)
            e, zz, sp = diag_rel(s.diag(), orc.diag(cur))
            assert e < 1e-12 and zz < 1e-12 and sp < 1e-12, (chunk, e, zz, sp)
        assert rel(s.get_state()[0], cur) < 1e-11


@pytest.mark.parametrize("batch", [1, 2])
def test_ns2d_large_grid_x_pass_forms(gpu, batch):
    # 1024^2, one member: 1024 x-rows, the split-row x-pass with 128-thread
    # teams (block barriers); two members: 2048 rows, the multi-row x-pass
    n = 1024
    g = O.grid2(n, n)
    z0 = np.stack([O.ic_ns2d_cfg(g, 4711 + m, 1.0) for m in range(batch)])
    orc = O.NS2D(n, v2=1 / 40, dt=0.005)
    with P.Stepper(flavour="ns2d", n=n, v2=1 / 40, dt=0.005, force="kolmogorov", kf=4.0, ampfor=-2.0,
                   batch=batch) as s:
        for m in range(batch):
            s.set_state(z0[m][None], m)
        s.run(2)
        for m in range(batch):
            assert rel(s.get_state(m)[0], orc.run(z0[m], 2)) < 1e-11, m


def test_ns2d_reference_ic(gpu):
    n = 128
    g = O.grid2(n, n)
    z0 = O.ic_ns2d_ref(g, 837920, 1.0)
    orc = O.NS2D(n, v2=1 / 40, dt=0.005)
    want = orc.run(z0, 40)
    with P.Stepper(flavour="ns2d", n=n, v2=1 / 40, dt=0.005, force="kolmogorov", kf=4.0, ampfor=-2.0) as s:
        s.set_state(z0[None])
        s.run(40)
        assert rel(s.get_state()[0], want) < 1e-11


def test_batch_members_independent(gpu):
    n = 32
    g = O.grid3(n, n, n)
    sts = [O.ic_solenoidal(g, 100 + m) for m in range(3)]
    orc = O.NS3D(n, v2=0.01, dt=0.002, force=3)
    with P.Stepper(flavour="ns3d", n=n, batch=3) as s:
        for m in range(3):
            s.set_state(sts[m], m)
        s.run(20)
        for m in range(3):
            assert rel(s.get_state(m), orc.run(sts[m], 20)) < 1e-12


def test_fp32_bulk_error_bound(gpu):
    n = 64
    st = O.ic_solenoidal(O.grid3(n, n, n), 23)
    orc = O.NS3D(n, v2=0.01, dt=0.002, force=3)
    want = orc.run(st, 100)
    with P.Stepper(flavour="ns3d", n=n, precision="fp32") as s:
        s.set_state(st)
        s.run(100)
        e = s.diag()[0]
    ew = orc.diag(want)[0]
    assert abs(e - ew) / ew < 1e-6


def test_timederiv_3d_and_2d(gpu):
    n = 32
    g = O.grid3(n, n, n)
    st = O.ic_solenoidal(g, 12)
    orc = O.NS3D(n, v2=0.02, dt=0.002, force=3)
    nu = O.nutab(g, 0.02, 0.002)[0]
    kx, ky, kz, L = O.tables(g)
    want = np.where(L == 1, -nu * st + orc.nonlin(st), 0)
    with P.Stepper(flavour="ns3d", n=n, v2=0.02) as s:
        assert rel(s.timederiv(st), want) < 1e-13
    g2 = O.grid2(64, 64)
    z = O.ic_ns2d_cfg(g2, 4)
    nu2 = O.nutab(g2, 1 / 40.0, 0.005)[0]
    kx, ky, _, L2 = O.tables(g2)
    want2 = np.where(L2 == 1, -nu2 * z + O.NS2D(64).nonlin(z), 0)
    with P.Stepper(flavour="ns2d", n=64, v2=1 / 40.0, dt=0.005, kf=4.0, ampfor=-2.0) as s:
        assert rel(s.timederiv(z[None])[0], want2) < 1e-13


@pytest.mark.parametrize("theta", [np.pi / 2, 90.0])
def test_bous_parity(gpu, theta):
    n = 32
    g = O.grid3(n, n, n)
    st = O.ic_solenoidal(g, 41, 0.5, 0.1)
    orc = O.BOUS(n, v2=1 / 100.0, sc=1.0, ri=1.0, theta=theta, dt=0.004)
    want = orc.run(st, 25)
    with P.Stepper(flavour="bous", n=n, v2=1 / 100.0, sc=1.0, ri=1.0, theta=theta, dt=0.004,
                   force="kolmogorov") as s:
        s.set_state(st)
        s.run(25)
        got = s.get_state()
        assert rel(s.timederiv(st), orc.rhs(st)) < 1e-13
    assert rel(got, want) < 1e-11
    assert rel(got[3], want[3]) < 1e-10


def test_bous_aspect_batch(gpu):
    g = O.grid3(64, 32, 32, 0.5)
    sts = [O.ic_solenoidal(g, 50 + m, 0.5, 0.1) for m in range(2)]
    orc = O.BOUS(32, v2=1 / 50.0, dt=0.004, alpha=0.5, nx=64, ny=32, nz=32)
    with P.Stepper(flavour="bous", nx=64, ny=32, nz=32, alpha=0.5, v2=1 / 50.0, dt=0.004, ri=1.0, batch=2) as s:
        for m in range(2):
            s.set_state(sts[m], m)
        s.run(10)
        for m in range(2):
            assert rel(s.get_state(m), orc.run(sts[m], 10)) < 1e-11


@pytest.mark.parametrize("variant", range(12))
def test_ns3d_xpass_variants_match_oracle(gpu, variant):
    n = 32
    st = O.ic_solenoidal(O.grid3(n, n, n), 24)
    want = O.NS3D(n, v2=0.01, dt=0.002, force=3).run(st, 20)
    with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, precision="fp64") as s:
        s.set_variant(variant)
        assert s.get_variant() == variant
        s.set_state(st)
        s.run(20)
        assert rel(s.get_state(), want) < 1e-11


@pytest.mark.parametrize("variant", range(12))
def test_bous_xpass_variants_match_oracle(gpu, variant):
    n = 32
    st = O.ic_solenoidal(O.grid3(n, n, n), 43, 0.5, 0.1)
    want = O.BOUS(n, v2=1 / 100.0, ri=1.0, theta=np.pi / 2, dt=0.004).run(st, 10)
    for prec, tol in (("fp64", 1e-11), ("fp32", 1e-4)):
        with P.Stepper(flavour="bous", n=n, v2=1 / 100.0, ri=1.0, theta=np.pi / 2, dt=0.004, precision=prec) as s:
            s.set_variant(variant)
            assert s.get_variant() == variant
            s.set_state(st)
            s.run(10)
            assert rel(s.get_state(), want) < tol, prec


@pytest.mark.parametrize("variant", range(2))
def test_ns3d_zfwd_variants_match_oracle(gpu, variant):
    n = 32
    st = O.ic_solenoidal(O.grid3(n, n, n), 27)
    want = O.NS3D(n, v2=0.01, dt=0.002, force=3).run(st, 20)
    with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, precision="fp64") as s:
        s.set_variant(variant, 1)
        assert s.get_variant(1) == variant
        s.set_state(st)
        s.run(20)
        assert rel(s.get_state(), want) < 1e-11


def test_autotune_keeps_state_and_picks_a_variant(gpu):
    n = 64
    st = O.ic_solenoidal(O.grid3(n, n, n), 25)
    with P.Stepper(flavour="ns3d", n=n) as s:
        s.set_state(st)
        rep = s.autotune(steps=2)
        assert "chosen xpass variant" in rep
        assert 0 <= s.get_variant() < 12
        assert np.array_equal(s.get_state(), st)
        s.run(10)
        got = s.get_state()
    assert rel(got, O.NS3D(n, v2=0.01, dt=0.002, force=3).run(st, 10)) < 1e-11


@pytest.mark.parametrize("flavour,kw", [
    ("ns3d", dict(n=32, v2=0.01, dt=0.002, force="kolmogorov")),
    ("euler", dict(n=32, dt=0.004, force="none")),
    ("ns2d", dict(n=64, v2=1 / 40.0, dt=0.005, force="kolmogorov", kf=4.0, ampfor=-2.0)),
    ("bous", dict(n=32, v2=0.01, dt=0.004, ri=1.0, force="kolmogorov")),
])
def test_graph_replays_match_direct_launches(gpu, flavour, kw):
    # captured step graphs replayed several times, with direct launches,
    # host copies, a dt change and synchronisation in between, give bitwise
    # the states of direct launches
    if flavour == "ns2d":
        st = O.ic_ns2d_cfg(O.grid2(64, 64), 6)[None]
    elif flavour == "bous":
        st = O.ic_solenoidal(O.grid3(32, 32, 32), 7, 0.5, 0.1)
    else:
        st = O.ic_solenoidal(O.grid3(32, 32, 32), 7)
    with P.Stepper(flavour=flavour, **kw) as a, P.Stepper(flavour=flavour, **kw) as b:
        a.set_graph(0)
        b.set_graph(4)
        a.set_state(st)
        b.set_state(st)
        for call, n in enumerate([13, 9, 4, 21]):
            if call == 2:
                a.set_dt(kw["dt"] * 0.5)
                b.set_dt(kw["dt"] * 0.5)
            a.run(n)
            b.run(n)
            sa, sb = a.get_state(), b.get_state()
            assert np.array_equal(sa, sb), (flavour, call)
            assert np.isfinite(sa).all()


def test_graph_tangent_linear_matches_direct(gpu):
    g = O.grid2(64, 64)
    z = O.ic_ns2d_cfg(g, 8)
    d = O.ic_ns2d_cfg(g, 9)
    with P.Stepper(flavour="ns2d", n=64, v2=1 / 40.0, dt=0.01, kf=4.0, ampfor=-2.0) as a, \
            P.Stepper(flavour="ns2d", n=64, v2=1 / 40.0, dt=0.01, kf=4.0, ampfor=-2.0) as b:
        a.set_graph(0)
        b.set_graph(5)
        for T in (0.37, 0.51):
            za, da = a.tangent(z[None], d[None], T, 0.01, 0)
            zb, db = b.tangent(z[None], d[None], T, 0.01, 0)
            assert np.array_equal(za, zb) and np.array_equal(da, db)


@pytest.mark.parametrize("variant", [0, 1, 5, 8, 9])
def test_fp32_xpass_variants_error_bound(gpu, variant):
    # FP32-bulk shapes (8, 9: radix-16 transforms) against the FP64 oracle
    n = 64
    st = O.ic_solenoidal(O.grid3(n, n, n), 26)
    orc = O.NS3D(n, v2=0.01, dt=0.002, force=3)
    want = orc.run(st, 30)
    with P.Stepper(flavour="ns3d", n=n, precision="fp32") as s:
        s.set_variant(variant)
        s.set_state(st)
        s.run(30)
        got = s.get_state()
        e = s.diag()[0]
    assert abs(e - orc.diag(want)[0]) / orc.diag(want)[0] < 1e-6
    assert rel(got, want) < 2e-5
