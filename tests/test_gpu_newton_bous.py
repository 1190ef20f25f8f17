"""The Boussinesq Newton-GMRES-hookstep with its two shift unknowns (x and z,
3D_Boussinesq/src/src_gmres/main.F90:767-778,883-908): the residual
S(sx, sz) Phi_T(z) - z and its normalisation ||(Phi_T(z), z)|| against the
oracle's flow map with the shift applied on the host; a few Newton steps from
a developed state re-checked on the oracle; and Newton steps that bring a
shifted guess back onto a steady state known in closed form.

The closed form: with no viscosity and no forcing, and gravity along z
(theta = pi/2), a y-directed velocity v(x, z) is an exact steady solution.
Its advection v d/dy vanishes on y-independent fields, the Lamb vector
u x omega = grad(v^2/2) has no curl, and u.g = 0 leaves rho = 0.  Modes with
kx = 1, 2 and kz = 0, 2, -1 make every shift other than (0, 0) change the
state, so a guess started at a shifted (sx, sz) must come back to (0, 0).
Every y-directed v(x, z) is steady too, so J has a null space (Newton may
move the state along that family) and the
inviscid linearisation puts its spectrum near the unit circle: GMRES meets
the Krylov cap and Newton closes in linearly, about 1.5x per step, instead
of converging to NETOL in a few steps."""
import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P

pytestmark = pytest.mark.gpu


def shift_host(g, z, sx, sz, sy=0.0):
    # S = exp(-i (kx sx + ky sy + kz sz)) on every field (main.F90:767-771)
    kx, ky, kz, L = O.tables(g)
    return z * np.exp(-1j * (kx * sx + ky * sy + kz * sz))[None, :]


def ner_host(g, z, zT, sx, sz):
    _, _, _, L = O.tables(g)
    F = (shift_host(g, zT, sx, sz) - z) * (L == 1)
    num = (np.abs(F) ** 2).sum()
    den = (np.abs(zT * (L == 1)) ** 2).sum() + (np.abs(z * (L == 1)) ** 2).sum()
    return np.sqrt(num / den)


def flow_host(n, kw, z, T, dt0):
    # the 3D time policy: INT(T/dt0) steps of dt0, then one of what is left
    # (3D_Boussinesq/src/src_gmres/main.F90:693-715)
    nst = int(T / dt0)
    out = O.BOUS(n, dt=dt0, **kw).run(z, nst) if nst > 0 else z.copy()
    last = T - nst * dt0
    return O.BOUS(n, dt=last, **kw).run(out, 1) if last > 0.0 else out


def steady_field(g):
    # vorticity of u = v(x, z) y-hat: omega = i k x (v_k y-hat) = (-i kz v_k, 0, i kx v_k)
    kx, ky, kz, L = O.tables(g)
    z = np.zeros((4, g.nkt), np.complex128)
    for (a, b, c) in [(1, 0, 1.0), (1, 2, 0.5 - 0.2j), (2, -1, 0.3j)]:
        ik = np.where((kx == a) & (ky == 0) & (kz == b))[0]
        assert len(ik) == 1 and L[ik[0]] == 1
        z[0, ik[0]] = -1j * b * c
        z[2, ik[0]] = 1j * a * c
    return z


@pytest.mark.parametrize("sx,sz", [(0.37, 1.1), (-0.8, 0.25)])
def test_bous_two_shift_residual_matches_oracle(gpu, sx, sz):
    n, v2, ri, theta, T, dt0 = 16, 0.02, 1.0, np.pi / 2, 0.1, 0.01
    g = O.grid3(n, n, n)
    z = O.BOUS(n, v2=v2, ri=ri, theta=theta, dt=0.01).run(O.ic_solenoidal(g, 71, 0.5, 0.2), 20)
    zT = flow_host(n, dict(v2=v2, ri=ri, theta=theta), z, T, dt0)
    want = ner_host(g, z, zT, sx, sz)
    with P.Stepper(flavour="bous", n=n, v2=v2, ri=ri, theta=theta, dt=dt0) as s:
        zr, res = s.newton(z, T=T, sx=sx, sz=sz, max_newton=0, dt0=dt0)
    print("NER gpu %.12e host %.12e" % (res["ner"], want), flush=True)
    assert abs(res["ner"] - want) <= 1e-10 * want
    assert np.array_equal(zr, z) and res["sx"] == sx and res["sz"] == sz
    # the z shift is applied with its own sign: flipping it changes the residual
    assert abs(ner_host(g, z, zT, sx, -sz) - want) > 1e-3 * want
    # the one-norm form (||Phi_T(z)|| alone) would differ by about sqrt(2)
    _, _, _, L = O.tables(g)
    one = np.sqrt((np.abs((shift_host(g, zT, sx, sz) - z) * (L == 1)) ** 2).sum() /
                  (np.abs(zT * (L == 1)) ** 2).sum())
    assert one / want > 1.3


@pytest.mark.parametrize("fixed_period", [1, 0])
def test_bous_newton_steps_recheck_on_oracle(gpu, fixed_period):
    # two Newton steps from a developed state with both shifts free: the returned
    # (z, T, sx, sz) must give the reported residual when re-integrated on the oracle
    n, v2, ri, theta, dt0 = 16, 0.02, 1.0, np.pi / 2, 0.01
    g = O.grid3(n, n, n)
    z0 = O.BOUS(n, v2=v2, ri=ri, theta=theta, dt=0.01).run(O.ic_solenoidal(g, 73, 0.5, 0.2), 30)
    with P.Stepper(flavour="bous", n=n, v2=v2, ri=ri, theta=theta, dt=dt0) as s:
        _, r0 = s.newton(z0, T=0.3, sx=0.05, sz=-0.05, max_newton=0, dt0=dt0)
        z, res = s.newton(z0, T=0.3, sx=0.05, sz=-0.05, max_newton=2, dt0=dt0, fixed_period=fixed_period,
                          max_gmres=30)
    print("NER %.6e -> %.6e after %d (T %.8f sx %.6f sz %.6f) status %s" % (
        r0["ner"], res["ner"], res["newton_iters"], res["T"], res["sx"], res["sz"], res["status_name"]), flush=True)
    assert res["newton_iters"] >= 1 and res["ner"] < r0["ner"] and res["T"] > 0.0
    if fixed_period:
        assert res["T"] == 0.3
    zT = flow_host(n, dict(v2=v2, ri=ri, theta=theta), z, res["T"], dt0)
    host = ner_host(g, z, zT, res["sx"], res["sz"])
    assert abs(host - res["ner"]) <= 1e-8 * res["ner"], (host, res["ner"])


def test_bous_newton_two_shifts_return_to_steady_state(gpu):
    n, ri, theta, T, dt0 = 16, 1.0, np.pi / 2, 0.5, 0.01
    g = O.grid3(n, n, n)
    zs = steady_field(g)
    orc = O.BOUS(n, v2=0.0, ri=ri, theta=theta, dt=dt0, ampfor=0.0)
    assert np.abs(orc.run(zs, 50) - zs).max() < 1e-12          # steady on the oracle too
    with P.Stepper(flavour="bous", n=n, v2=0.0, ri=ri, theta=theta, dt=dt0, ampfor=0.0) as s:
        _, r0 = s.newton(zs, T=T, sx=0.06, sz=-0.04, max_newton=0, dt0=dt0)
        z, res = s.newton(zs, T=T, sx=0.06, sz=-0.04, fixed_period=1, dt0=dt0, max_gmres=40, max_newton=20)
    print("NER %.3e -> %.3e: %s after %d Newton, %d GMRES; sx %.3e sz %.3e" % (
        r0["ner"], res["ner"], res["status_name"], res["newton_iters"], res["gmres_iters"], res["sx"], res["sz"]),
        flush=True)
    assert r0["ner"] > 1e-3
    assert res["ner"] < 1e-3 * r0["ner"], res
    assert abs(res["sx"]) < 1e-5 and abs(res["sz"]) < 1e-5
    # the state may move along the family of steady states, but stays in it
    zT = orc.run(z, int(T / dt0))
    host = ner_host(g, z, zT, res["sx"], res["sz"])
    assert abs(host - res["ner"]) <= 1e-6 * res["ner"], (host, res["ner"])
    assert ner_host(g, z, zT, 0.0, 0.0) < 1e-3 * r0["ner"]
    assert np.abs(z).max() > 0.5 * np.abs(zs).max()
