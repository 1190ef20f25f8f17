"""Floquet analysis (psg_floquet) on the GPU.

- Closed forms: at the rest state of the unforced flow the monodromy operator
  is diagonal, each mode multiplied per step by the Crank-Nicolson factor
  (1 - a)/(1 + a) (2D: a = dt nu |k|^2 / 2; 3D: the NUZN NU1 table product),
  so the leading multipliers are those of the smallest shells, raised to the
  number of steps.
- The laminar Kolmogorov flow at Re = 40 (a fixed point, unstable): each
  returned Ritz pair is checked against the oracle's flow map, A x - mu x with
  A x from central differences of the oracle (the reference's own eigenvector
  check, 2D_NavierStokes/src_RK4/src_gmres/main.F90:973-992), and the
  tangent-linear and finite-difference Jv give the same multipliers.
Reference: ARPACK dnaupd loop, 2D_NavierStokes/src_RK4/src_gmres/main.F90:859-1037.
"""
import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P

pytestmark = pytest.mark.gpu


def ip(a, b, L):
    return float(np.sum((np.conj(a) * b).real * L))


def test_floquet_2d_rest_state_closed_form(gpu):
    n, v2, T, dt0 = 64, 1 / 40.0, 0.5, 0.005
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    nst = int(T / dt0)
    dt = T / nst
    k2 = np.unique(np.round((kx ** 2 + ky ** 2)[L == 1], 9))
    mult = ((1 - 0.5 * dt * v2 * k2) * (1.0 / (1 + 0.5 * dt * v2 * k2))) ** nst
    want = np.sort(mult)[::-1][:4]
    with P.Stepper(flavour="ns2d", n=n, v2=v2, dt=dt0, force="none") as s:
        for jv in (1, 0):
            # finite differences carry ~eps relative noise in every Jv: their Ritz
            # residuals level off there, so they are asked for 1e-5 only
            mu, err, _, res = s.floquet(np.zeros((1, g.nkt), complex), T, nev=4, ncv=30, jv=jv, dt0=dt0,
                                        tol=1e-7 if jv else 1e-5)
            print("2D rest state jv=%d:" % jv, mu, err, res, flush=True)
            assert np.abs(mu.imag).max() < 1e-12
            tol = 1e-12 if jv else 1e-5
            assert np.allclose(mu.real, want, rtol=tol, atol=0), (mu, want)
            assert res["nconv"] == 4


def test_floquet_3d_rest_state_closed_form(gpu):
    # nu T large enough that the leading shells' multipliers are well apart
    n, v2, dt0 = 32, 0.05, 0.004
    T = 0.5 * (1 + 1e-12)
    nst = int(T / dt0)                                      # the uniform time policy: dt = T / INT(T/dt0)
    g = O.grid3(n, n, n)
    kx, ky, kz, L = O.tables(g)
    _, nuzn, nu1 = O.nutab(g, v2, T / nst)
    fac = (nuzn * nu1)[L == 1]
    k2 = (kx ** 2 + ky ** 2 + kz ** 2)[L == 1]
    shells = {}
    for a, b in zip(np.round(k2, 9), fac):
        shells.setdefault(a, b)
    want = np.sort(np.array(list(shells.values())) ** nst)[::-1][:3]
    with P.Stepper(flavour="ns3d", n=n, v2=v2, dt=dt0, force="none") as s:
        mu, err, _, res = s.floquet(np.zeros((3, g.nkt), complex), T, nev=3, ncv=30, jv=1, dt0=dt0)
    print("3D rest state:", mu, want, res, flush=True)
    assert np.allclose(mu.real, want, rtol=1e-12, atol=0) and np.abs(mu.imag).max() < 1e-12
    assert res["nconv"] == 3


def test_floquet_kolmogorov_laminar_eigenpairs_vs_oracle(gpu):
    n, v2, T, dt0 = 64, 1 / 40.0, 1.0, 0.005
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    z = np.zeros(g.nkt, complex)
    z[(kx == 0) & (ky == 4)] = -2.0 / (16 * v2)           # the laminar fixed point F / (nu k^2)
    orc = O.NS2D(n, v2=v2, dt=dt0)
    assert np.abs(orc.run(z, 20) - z).max() < 1e-13       # it is one for the oracle's scheme too
    nst = int(T / dt0 + 1e-9)
    with P.Stepper(flavour="ns2d", n=n, v2=v2, dt=dt0, force="kolmogorov", kf=4.0, ampfor=-2.0) as s:
        mu, err, vec, res = s.floquet(z[None], T, nev=6, ncv=60, jv=1, dt0=dt0, vectors=True)
        mu_fd, _, _, res_fd = s.floquet(z[None], T, nev=6, ncv=60, jv=0, dt0=dt0)
    print("laminar Kolmogorov multipliers (TL):", mu, err, res, flush=True)
    print("laminar Kolmogorov multipliers (FD):", mu_fd, res_fd, flush=True)
    assert res["nconv"] == 6
    assert abs(mu[0]) > 1.0                                 # the laminar flow is unstable at Re = 40
    assert np.abs(mu - mu_fd).max() <= 1e-5 * abs(mu[0])

    def A(x):                                               # oracle central difference of Phi_T
        h = 1e-5 * np.sqrt(ip(z, z, L)) / np.sqrt(ip(x, x, L))
        return (orc.run(z + h * x, nst) - orc.run(z - h * x, nst)) / (2 * h)

    i = 0
    while i < len(mu):
        if abs(mu[i].imag) > 1e-12 * abs(mu[i]) and i + 1 < len(mu):
            xr, xi = vec[i, 0], vec[i + 1, 0]
            a, b = mu[i].real, mu[i].imag
            r1 = A(xr) - (a * xr - b * xi)
            r2 = A(xi) - (b * xr + a * xi)
            rel = np.sqrt(ip(r1, r1, L) + ip(r2, r2, L)) / (abs(mu[i]) * np.sqrt(ip(xr, xr, L) + ip(xi, xi, L)))
            i += 2
        else:
            x = vec[i, 0]
            r = A(x) - mu[i].real * x
            rel = np.sqrt(ip(r, r, L)) / (abs(mu[i]) * np.sqrt(ip(x, x, L)))
            i += 1
        print("eigenpair %d: |A x - mu x| / |mu x| = %.2e" % (i, rel), flush=True)
        assert rel < 1e-6
