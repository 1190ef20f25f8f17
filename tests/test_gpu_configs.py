"""GPU parity at the benchmark configurations' own sizes: the kernel instances
the benchmark times are the ones checked here against the FP64 CPU oracle.

- 3D Navier-Stokes 256^3, nu = 1/100, dt = 0.002, Kolmogorov forcing, the
  benchmark's seeded solenoidal IC (seed 20240601): 100 Heun+CN steps of the
  plain FP64 form against the oracle (energy, enstrophy and shell spectrum
  <= 1e-12 relative every 25 steps, state <= 1e-11), every x-pass and
  update-pass shape variant for 2 steps, and the FP32-bulk form against the
  same oracle trajectory (energy <= 1e-6 relative).
- 2D Navier-Stokes 512^2, Re = 40 (the Newton grid): 100 RK4+CN steps against
  the oracle at the same bars, and the tangent linear at 512^2 against central
  differences of the oracle's flow map.
- 2D 256^2 (the 2D stepping configuration): 1000 steps against the oracle.
- 3D 128^3 over 200 steps with checkpoints every 20 (the adjoint
  configuration): the dot-product identity of tangent linear and adjoint, and
  the gradient of J = ||Phi_T(u0) - u0||^2 against central differences of the
  oracle's J (epsilon 1e-6 and a fourth-order difference at h = 1e-4).
- 512-point pencils: boxes with one extent of 512 run the x-, y- and z-pass
  instances of the 512^3 Boussinesq ensemble, Navier-Stokes and Boussinesq,
  against the oracle.
- 3D Boussinesq 512^3 x 4 members at full size: size-independent properties
  (each member bitwise equal to a single-member run of the same state, the
  vorticity divergence-free after the Leray step, finite energies).

Reference: 3D_NavierStokes/src/cuda.cu:60-112 (the time loop),
2D_NavierStokes/src_RK4/src_gmres/cuda.cu:44-76, 3D_Boussinesq/src/cuda.cu:65-101.
"""
import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P

pytestmark = pytest.mark.gpu

SEED_256 = 20240601


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def diag_rel(dg, do):
    return abs(dg[0] - do[0]) / abs(do[0]), abs(dg[1] - do[1]) / abs(do[1]), rel(dg[2], do[2])


def ip(a, b, L):
    return float(np.sum((np.conj(a) * b).real * L[None, :]))


# ---------------------------------------------------------------- 3D 256^3
@pytest.fixture(scope="module")
def ns3d_256():
    """Oracle trajectory of the 256^3 workload: states after 0, 2, 25, 50, 75, 100 steps."""
    n = 256
    g = O.grid3(n, n, n)
    st = O.ic_solenoidal(g, SEED_256, 0.5)
    orc = O.NS3D(n, v2=0.01, dt=0.002, force=3)
    traj = {0: st, 2: orc.run(st, 2)}
    cur = st
    for k in (25, 50, 75, 100):
        cur = orc.run(cur, 25)
        traj[k] = cur
    print("oracle 256^3 trajectory ready", flush=True)
    return orc, traj


def test_config2_ns3d_256_fp64_100_steps(gpu, ns3d_256):
    orc, traj = ns3d_256
    with P.Stepper(flavour="ns3d", n=256, v2=0.01, dt=0.002, force="kolmogorov") as s:
        s.set_state(traj[0])
        rep = s.autotune(steps=2)          # the benchmark's own shape selection
        print(rep, flush=True)
        assert np.array_equal(s.get_state(), traj[0])
        for k in (25, 50, 75, 100):
            s.run(25)
            e, z, sp = diag_rel(s.diag(), orc.diag(traj[k]))
            print("256^3 fp64 step %d: energy %.2e enstrophy %.2e spectrum %.2e" % (k, e, z, sp), flush=True)
            assert e < 1e-12 and z < 1e-12 and sp < 1e-12, (k, e, z, sp)
        assert rel(s.get_state(), traj[100]) < 1e-11


@pytest.mark.parametrize("prec,variant,zvar", [("fp64", v, 0) for v in range(12)] + [("fp64", 1, 1)] +
                         [("fp32", v, 0) for v in range(12)] + [("fp32", 1, 1)])
def test_config2_ns3d_256_pass_variants(gpu, ns3d_256, prec, variant, zvar):
    orc, traj = ns3d_256
    with P.Stepper(flavour="ns3d", n=256, v2=0.01, dt=0.002, force="kolmogorov", precision=prec) as s:
        s.set_variant(variant, 0)
        s.set_variant(zvar, 1)
        s.set_state(traj[0])
        s.run(2)
        err = rel(s.get_state(), traj[2])
    assert err < (1e-12 if prec == "fp64" else 2e-5), err


def test_config2_ns3d_256_fp32_bulk_vs_oracle(gpu, ns3d_256):
    orc, traj = ns3d_256
    with P.Stepper(flavour="ns3d", n=256, v2=0.01, dt=0.002, force="kolmogorov", precision="fp32") as s:
        s.set_state(traj[0])
        s.autotune(steps=2)
        errs = []
        for k in (25, 50, 75, 100):
            s.run(25)
            e = s.diag()[0]
            eo = orc.diag(traj[k])[0]
            errs.append(abs(e - eo) / eo)
        print("256^3 fp32-bulk energy error vs oracle every 25 steps:", errs, flush=True)
        assert max(errs) < 1e-6, errs
        assert rel(s.get_state(), traj[100]) < 1e-4


# ---------------------------------------------------------------- 2D
def test_config4_ns2d_512_100_steps(gpu):
    n = 512
    g = O.grid2(n, n)
    z0 = O.ic_ns2d_cfg(g, 20240607, 1.0)
    orc = O.NS2D(n, v2=1 / 40, dt=0.005)
    with P.Stepper(flavour="ns2d", n=n, v2=1 / 40, dt=0.005, force="kolmogorov", kf=4.0, ampfor=-2.0) as s:
        s.set_state(z0[None])
        cur = z0
        for k in range(4):
            s.run(25)
            cur = orc.run(cur, 25)
            e, zz, sp = diag_rel(s.diag(), orc.diag(cur))
            assert e < 1e-12 and zz < 1e-12 and sp < 1e-12, (k, e, zz, sp)
        assert rel(s.get_state()[0], cur) < 1e-11


def test_config4_tangent_linear_512_vs_oracle_fd(gpu):
    n, T, dt0, v2 = 512, 0.25, 0.005, 1 / 40.0
    g = O.grid2(n, n)
    z = O.NS2D(n, v2=v2, dt=dt0).run(O.ic_ns2d_cfg(g, 20240607, 1.0), 40)
    _, _, _, L = O.tables(g)
    rng = np.random.default_rng(5)
    d = (rng.standard_normal(g.nkt) + 1j * rng.standard_normal(g.nkt)) * (L == 1)
    d /= np.sqrt((np.abs(d) ** 2).sum())
    nst = int(T / dt0 + 1e-9)
    orc = O.NS2D(n, v2=v2, dt=T / nst)
    eps = 1e-6
    fd = (orc.run(z + eps * d, nst) - orc.run(z - eps * d, nst)) / (2 * eps)
    for jv_prec, tol in (("fp64", 1e-7), ("fp32", 1e-4)):
        with P.Stepper(flavour="ns2d", n=n, v2=v2, dt=dt0, kf=4.0, ampfor=-2.0, precision=jv_prec) as s:
            z1, dz1 = s.tangent(z[None], d[None], T, dt0, 0)
        err = rel(dz1[0], fd)
        print("512^2 tangent linear %s vs oracle central difference: %.2e" % (jv_prec, err), flush=True)
        assert err < tol, (jv_prec, err)


def test_config1_ns2d_256_1000_steps(gpu):
    n = 256
    g = O.grid2(n, n)
    z0 = O.ic_ns2d_cfg(g, 837920, 1.0)
    orc = O.NS2D(n, v2=1 / 40, dt=0.005)
    worst = 0.0
    with P.Stepper(flavour="ns2d", n=n, v2=1 / 40, dt=0.005, force="kolmogorov", kf=4.0, ampfor=-2.0) as s:
        s.set_state(z0[None])
        cur = z0
        for k in range(10):
            s.run(100)
            cur = orc.run(cur, 100)
            e, zz, sp = diag_rel(s.diag(), orc.diag(cur))
            worst = max(worst, e, zz, sp)
            if k == 0:
                assert e < 1e-12 and zz < 1e-12 and sp < 1e-12, (e, zz, sp)
        print("256^2 1000 steps: worst diagnostic error %.2e" % worst, flush=True)
        assert worst < 1e-11
        assert rel(s.get_state()[0], cur) < 1e-10


# ---------------------------------------------------------------- 3D 128^3 adjoint
N3, STEPS3, CKPT3 = 128, 200, 20


@pytest.fixture(scope="module")
def adj128():
    g = O.grid3(N3, N3, N3)
    L = O.tables(g)[3].astype(float)
    z0 = O.ic_solenoidal(g, 20240603, 0.5)
    orc = O.NS3D(N3, v2=0.01, dt=0.002, force=3)

    def J(x):
        r = (orc.run(x, STEPS3) - x) * L[None, :]
        return float(np.sum(np.abs(r) ** 2))

    with P.Stepper(flavour="ns3d", n=N3, v2=0.01, dt=0.002, force="kolmogorov") as s:
        Jg, grad, zT = s.residual_grad(z0, STEPS3, CKPT3)
    with P.Stepper(flavour="ns3d", n=N3, v2=0.01, dt=0.002, force="kolmogorov", precision="fp32") as s:
        _, grad32, _ = s.residual_grad(z0, STEPS3, CKPT3)
    return g, L, z0, orc, J, Jg, grad, grad32, zT


def direction(g, seed):
    v = O.ic_solenoidal(g, seed, 0.5)
    return v / np.sqrt(np.sum(np.abs(v) ** 2))


def test_config3_forward_map_and_J_vs_oracle(gpu, adj128):
    g, L, z0, orc, J, Jg, grad, grad32, zT = adj128
    want = orc.run(z0, STEPS3)
    assert rel(zT, want) < 1e-11
    J0 = float(np.sum(np.abs((want - z0) * L[None, :]) ** 2))
    assert abs(Jg - J0) <= 1e-11 * J0


def test_config3_dot_product_identity(gpu, adj128):
    g, L, z0, *_ = adj128
    v, w = direction(g, 777100), direction(g, 777101)
    with P.Stepper(flavour="ns3d", n=N3, v2=0.01, dt=0.002, force="kolmogorov") as s:
        _, jv = s.tangent_steps(z0, v, STEPS3)
        _, jtw = s.adjoint(z0, w, STEPS3, CKPT3)
    lhs, rhs = ip(jv, w, L), ip(v, jtw, L)
    print("128^3 dot product: %.16e %.16e" % (lhs, rhs), flush=True)
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)


def test_config3_gradient_central_difference_eps_1e6(gpu, adj128):
    # the configuration's literal check: (J(u0+eps v) - J(u0-eps v)) / (2 eps), eps = 1e-6;
    # its round-off is ~1e-16 J / eps, so the error is also stated over |grad|
    g, L, z0, orc, J, Jg, grad, grad32, _ = adj128
    eps = 1e-6
    v = direction(g, 777000)
    fd = (J(z0 + eps * v) - J(z0 - eps * v)) / (2 * eps)
    ad, ad32 = ip(grad, v, L), ip(grad32, v, L)
    gn = np.sqrt(ip(grad, grad, L))
    print("128^3 grad.v %.12e fd %.12e rel %.2e over |grad| %.2e fp32 rel %.2e" %
          (ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / gn, abs(ad32 - fd) / abs(fd)), flush=True)
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    assert abs(ad32 - fd) <= 1e-4 * abs(fd), (ad32, fd)


def test_config3_gradient_fourth_order_difference(gpu, adj128):
    g, L, z0, orc, J, Jg, grad, grad32, _ = adj128
    h = 1e-4
    v = direction(g, 777001)
    fd = (8 * (J(z0 + h * v) - J(z0 - h * v)) - (J(z0 + 2 * h * v) - J(z0 - 2 * h * v))) / (12 * h)
    ad = ip(grad, v, L)
    print("128^3 grad.v %.12e fourth-order fd %.12e rel %.2e" % (ad, fd, abs(ad - fd) / abs(fd)), flush=True)
    assert abs(ad - fd) <= 1e-7 * abs(fd), (ad, fd)


# ---------------------------------------------------------------- 512-point pencils
@pytest.mark.parametrize("box", [(512, 32, 32, 1 / 16.0), (32, 512, 32, 1.0), (32, 32, 512, 1.0)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pencils_512_ns3d(gpu, box, prec):
    nx, ny, nz, alpha = box
    g = O.grid3(nx, ny, nz, alpha)
    st = O.ic_solenoidal(g, 512 + nx + 2 * ny, 0.5)
    orc = O.NS3D(ny, v2=0.01, dt=0.002, force=3, alpha=alpha, nx=nx, ny=ny, nz=nz)
    want = orc.run(st, 4)
    with P.Stepper(flavour="ns3d", nx=nx, ny=ny, nz=nz, alpha=alpha, v2=0.01, dt=0.002, force="kolmogorov",
                   precision=prec) as s:
        assert rel(s.nonlin(st), orc.nonlin(st)) < (1e-13 if prec == "fp64" else 1e-5)
        s.set_state(st)
        s.run(4)
        err = rel(s.get_state(), want)
    assert err < (1e-12 if prec == "fp64" else 2e-5), err


@pytest.mark.parametrize("box", [(512, 32, 32, 1 / 16.0), (32, 512, 32, 1.0), (32, 32, 512, 1.0)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pencils_512_bous(gpu, box, prec):
    nx, ny, nz, alpha = box
    g = O.grid3(nx, ny, nz, alpha)
    st = O.ic_solenoidal(g, 4096 + nx + 2 * ny, 0.5, 0.1)
    orc = O.BOUS(ny, v2=1 / 400.0, sc=1.0, ri=1.0, theta=90.0, dt=0.002, alpha=alpha, nx=nx, ny=ny, nz=nz)
    want = orc.run(st, 3)
    with P.Stepper(flavour="bous", nx=nx, ny=ny, nz=nz, alpha=alpha, v2=1 / 400.0, sc=1.0, ri=1.0, theta=90.0,
                   dt=0.002, force="kolmogorov", precision=prec) as s:
        assert rel(s.timederiv(st), orc.rhs(st)) < (1e-13 if prec == "fp64" else 1e-5)
        s.set_state(st)
        s.run(3)
        err = rel(s.get_state(), want)
    assert err < (1e-11 if prec == "fp64" else 2e-5), err


# ---------------------------------------------------------------- Boussinesq 512^3 x 4
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_config5_bous_512_ensemble_properties(gpu, prec):
    n, members, steps = 512, 4, 2
    g = P.hgrid("bous", n, n, n)
    kx, ky, kz, L = P.tables(g)
    sts = [P.ic_solenoidal(g, 20240605 + m, 0.5, 0.1) for m in range(members)]
    kw = dict(flavour="bous", n=n, v2=1 / 400.0, sc=1.0, ri=1.0, theta=90.0, dt=0.002, force="kolmogorov",
              precision=prec)
    with P.Stepper(batch=members, **kw) as s:
        s.set_graph(2)
        for m in range(members):
            s.set_state(sts[m], m)
        s.run(steps)
        got = [s.get_state(m) for m in range(members)]
        en = [s.diag(m)[0] for m in range(members)]
    with P.Stepper(batch=1, **kw) as s1:
        s1.set_graph(2)
        s1.set_state(sts[2])
        s1.run(steps)
        single = s1.get_state()
    assert np.array_equal(got[2], single)
    for m in range(members):
        w = got[m]
        assert np.isfinite(w).all() and np.isfinite(en[m]) and en[m] > 0
        div = kx * w[0] + ky * w[1] + kz * w[2]
        mag = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2) * np.sqrt(np.abs(w[0]) ** 2 + np.abs(w[1]) ** 2 +
                                                              np.abs(w[2]) ** 2)
        assert np.abs(div).max() <= 1e-12 * mag.max(), m
        assert not np.array_equal(got[m], sts[m])
