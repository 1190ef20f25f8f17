"""GPU tests of the tangent-linear stepper and the Newton-GMRES-hookstep
solver.  The tangent linear is checked against central finite differences of
the FP64 oracle's flow map; Newton against fixed points whose location is
known in closed form."""
import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P

pytestmark = pytest.mark.gpu


def masked_random(g, seed, scale=1.0):
    kx, ky, kz, L = O.tables(g)
    rng = np.random.default_rng(seed)
    d = (rng.standard_normal(g.nkt) + 1j * rng.standard_normal(g.nkt)) * (L == 1)
    return scale * d / np.sqrt((np.abs(d) ** 2).sum())


def test_tl_2d_matches_oracle_central_fd(gpu):
    n, T, dt0, v2 = 64, 0.4, 0.01, 1 / 40.0
    g = O.grid2(n, n)
    z = O.NS2D(n, v2=v2, dt=dt0).run(O.ic_ns2d_cfg(g, 3, 1.0), 50)   # a developed state
    d = masked_random(g, 11)
    nst = int(T / dt0)
    orc = O.NS2D(n, v2=v2, dt=T / nst)
    eps = 1e-6
    fd = (orc.run(z + eps * d, nst) - orc.run(z - eps * d, nst)) / (2 * eps)
    with P.Stepper(flavour="ns2d", n=n, v2=v2, dt=dt0, kf=4.0, ampfor=-2.0) as s:
        z1, dz1 = s.tangent(z[None], d[None], T, dt0, 0)
    assert np.abs(z1[0] - orc.run(z, nst)).max() < 1e-11 * np.abs(z1).max()
    err = np.abs(dz1[0] - fd).max() / np.abs(fd).max()
    assert err < 1e-7, err
    # linearity in the direction
    with P.Stepper(flavour="ns2d", n=n, v2=v2, dt=dt0, kf=4.0, ampfor=-2.0) as s:
        _, dz2 = s.tangent(z[None], 3.0 * d[None], T, dt0, 0)
    assert np.abs(dz2 - 3.0 * dz1).max() < 1e-12 * np.abs(dz2).max()


@pytest.mark.parametrize("policy", [0, 1])
def test_tangent_about_recorded_base(gpu, policy):
    # psg_tangent_many records the base flow's stages once and integrates each
    # direction alone against the record: the same Jv as the combined stepper,
    # and the oracle's central difference
    n, v2, dt0 = 64, 1 / 40.0, 0.01
    T = 0.4 if policy == 0 else 0.437
    g = O.grid2(n, n)
    z = O.NS2D(n, v2=v2, dt=dt0).run(O.ic_ns2d_cfg(g, 3, 1.0), 50)
    ds = np.array([masked_random(g, 30 + i) for i in range(3)])
    with P.Stepper(flavour="ns2d", n=n, v2=v2, dt=dt0, kf=4.0, ampfor=-2.0) as s:
        z1, dm = s.tangent_many(z[None], ds[:, None, :], T, dt0, policy)
        _, d32 = s.tangent_many(z[None], ds[:, None, :], T, dt0, policy, fp32=True)
        for i in range(3):
            zc, dc = s.tangent(z[None], ds[i][None], T, dt0, policy)
            assert np.abs(dm[i] - dc).max() <= 1e-12 * np.abs(dc).max(), i
            assert np.abs(z1 - zc).max() <= 1e-13 * np.abs(zc).max()
            err32 = np.abs(d32[i] - dc).max() / np.abs(dc).max()
            assert 0 < err32 < 1e-5, err32
    # an FP32-bulk context records in FP32 and replays in FP32 passes
    with P.Stepper(flavour="ns2d", n=n, v2=v2, dt=dt0, kf=4.0, ampfor=-2.0, precision="fp32") as s32:
        _, db = s32.tangent_many(z[None], ds[:, None, :], T, dt0, policy)
        for i in range(3):
            _, dc = s32.tangent(z[None], ds[i][None], T, dt0, policy)
            err = np.abs(db[i] - dc).max() / np.abs(dc).max()
            assert err < 1e-5, err
    nst = int(T / dt0)
    eps = 1e-6
    if policy == 0:
        orc = O.NS2D(n, v2=v2, dt=T / nst)
        fd = (orc.run(z + eps * ds[0], nst) - orc.run(z - eps * ds[0], nst)) / (2 * eps)
        assert np.abs(dm[0, 0] - fd).max() / np.abs(fd).max() < 1e-7


@pytest.mark.parametrize("jv", [1, 2])
def test_newton_recorded_and_combined_tangent_agree(gpu, jv):
    # the tangent-linear Jv against the recorded base takes the solver through
    # the same iterations as the combined base + perturbation stepper
    n, re_ = 32, 5.0
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    lam = np.zeros(g.nkt, complex)
    lam[np.where((kx == 0) & (ky == 4))[0][0]] = -re_ / 8
    guess = lam + masked_random(g, 7, 0.05)
    out = []
    for rec in (0, 1):
        with P.Stepper(flavour="ns2d", n=n, v2=1 / re_, dt=0.01, kf=4.0, ampfor=-2.0) as s:
            out.append(s.newton(guess[None], T=2.0, nshift=0, fixed_period=1, jv=jv, dt0=0.01, max_gmres=40,
                                tl_record=rec))
    (z0, r0), (z1, r1) = out
    print(r0, r1, flush=True)
    assert r0["status_name"] == "CONVERGED" == r1["status_name"]
    assert r0["newton_iters"] == r1["newton_iters"] and r0["gmres_iters"] == r1["gmres_iters"]
    assert np.abs(z1 - z0).max() < 1e-12 and np.abs(z1[0] - lam).max() < 1e-10


@pytest.mark.parametrize("jv", [0, 1, 2])
def test_newton_converges_to_laminar_equilibrium(gpu, jv):
    # below the first instability the Kolmogorov laminar state is the unique
    # nearby fixed point of Phi_T: omega(0,4) = -Re/8 (2D forcing -2 at (0,4))
    n, re_ = 32, 5.0
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    lam = np.zeros(g.nkt, complex)
    i4 = np.where((kx == 0) & (ky == 4))[0][0]
    lam[i4] = -re_ / 8
    guess = lam + masked_random(g, 7, 0.05)
    with P.Stepper(flavour="ns2d", n=n, v2=1 / re_, dt=0.01, kf=4.0, ampfor=-2.0) as s:
        z, res = s.newton(guess[None], T=2.0, nshift=0, fixed_period=1, jv=jv, dt0=0.01, max_gmres=40)
    assert res["status_name"] == "CONVERGED", res
    assert res["ner"] <= 1e-12
    assert np.abs(z[0] - lam).max() < 1e-10


def test_newton_converged_orbit_rechecks(gpu):
    # the reference's own consistency check (2D_NavierStokes/src_RK4/main.F90:249-291):
    # re-integrating a converged solution gives the same residual
    n, re_ = 32, 5.0
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    lam = np.zeros(g.nkt, complex)
    lam[np.where((kx == 0) & (ky == 4))[0][0]] = -re_ / 8
    with P.Stepper(flavour="ns2d", n=n, v2=1 / re_, dt=0.01, kf=4.0, ampfor=-2.0) as s:
        z, res = s.newton((lam + masked_random(g, 9, 0.02))[None], T=1.5, nshift=0, fixed_period=1, dt0=0.01)
        zT = s.flow_map(z, res["T"], 0.01, 0)
    ner = np.sqrt((np.abs(zT - z) ** 2).sum()) / np.sqrt((np.abs(zT) ** 2).sum())
    assert abs(ner - res["ner"]) <= 1e-3 * max(res["ner"], 1e-16) + 1e-15


def test_newton_ner2_low_wavenumber_residual(gpu):
    # NER2 (2D_NavierStokes/src_RK4/src_gmres/main.F90:124-125, 737-751): the
    # residual over the modes with |k|^2 <= KYmax2 = (IKTY-KTY-1)^2/4, relative
    # to ||Phi_T(z)||; recomputed on the host from the state the solver returns
    n, re_ = 32, 5.0
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    lam = np.zeros(g.nkt, complex)
    lam[np.where((kx == 0) & (ky == 4))[0][0]] = -re_ / 8
    with P.Stepper(flavour="ns2d", n=n, v2=1 / re_, dt=0.01, kf=4.0, ampfor=-2.0) as s:
        z, res = s.newton((lam + masked_random(g, 11, 0.2))[None], T=1.5, nshift=0, fixed_period=1, dt0=0.01,
                          max_newton=1)
        zT = s.flow_map(z, res["T"], 0.01, 0)[0]
    F = (zT - z[0]) * (L == 1)
    low = (kx ** 2 + ky ** 2) <= g.kty ** 2 / 4.0
    ner = np.sqrt((np.abs(F) ** 2).sum()) / np.sqrt((np.abs(zT) ** 2).sum())
    ner2 = np.sqrt((np.abs(F[low]) ** 2).sum()) / np.sqrt((np.abs(zT) ** 2).sum())
    print("NER %.6e (host %.6e)  NER2 %.6e (host %.6e)" % (res["ner"], ner, res["ner2"], ner2), flush=True)
    assert 0.0 < res["ner2"] <= res["ner"]
    assert abs(res["ner2"] - ner2) <= 1e-9 * ner2 and abs(res["ner"] - ner) <= 1e-9 * ner


def test_newton_batch_matches_one_guess_at_a_time(gpu):
    # five guesses of different periods on three lanes: every guess must reach the
    # laminar fixed point exactly as the one-guess solver does from the same start
    n, re_ = 32, 5.0
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    lam = np.zeros(g.nkt, complex)
    lam[np.where((kx == 0) & (ky == 4))[0][0]] = -re_ / 8
    guesses = np.array([lam + masked_random(g, 20 + i, 0.05) for i in range(5)])[:, None, :]
    Ts = np.array([1.5, 2.0, 2.5, 1.75, 3.0])
    opts = dict(nshift=0, fixed_period=1, jv=0, dt0=0.01, max_gmres=40)
    with P.Stepper(flavour="ns2d", n=n, v2=1 / re_, dt=0.01, kf=4.0, ampfor=-2.0, batch=3) as sb:
        zb, Tb, _, rb = sb.newton_batch(guesses, Ts, **opts)
    with P.Stepper(flavour="ns2d", n=n, v2=1 / re_, dt=0.01, kf=4.0, ampfor=-2.0) as s1:
        for i in range(5):
            z1, r1 = s1.newton(guesses[i], T=Ts[i], **opts)
            assert rb[i]["status_name"] == "CONVERGED" == r1["status_name"], (i, rb[i], r1)
            assert rb[i]["newton_iters"] == r1["newton_iters"] and rb[i]["gmres_iters"] == r1["gmres_iters"]
            assert abs(rb[i]["ner"] - r1["ner"]) <= 1e-3 * r1["ner"] + 1e-15
            assert np.abs(zb[i, 0] - z1[0]).max() < 1e-11
            assert np.abs(zb[i, 0] - lam).max() < 1e-10
            assert Tb[i] == Ts[i]


def test_newton_batch_rejects_what_it_does_not_batch(gpu):
    with P.Stepper(flavour="ns2d", n=32, v2=0.2, dt=0.01, batch=2) as sb:
        g = np.zeros((2, 1, sb.nkt), complex)
        with pytest.raises(P.PsgError, match="finite-difference"):
            sb.newton_batch(g, [1.0, 1.0], jv=1)
    with P.Stepper(flavour="ns3d", n=16, batch=2) as s3:
        with pytest.raises(P.PsgError, match="2D"):
            s3.newton_batch(np.zeros((2, 3, s3.nkt), complex), [1.0, 1.0])
