"""CPU pins of the oracle's diagnostics, forcing modes and configuration
initial condition, each against something the oracle does not compute itself:

- or3_diag / or2_diag energy and enstrophy against Parseval: the physical
  velocity and vorticity are synthesised here with numpy's inverse FFT from
  the half-space state (Biot-Savart written out in numpy), and the mean
  squares compared with the spectral sums (E = 1/2 <|u|^2>, 3D enstrophy
  1/2 <|w|^2>, 2D enstrophy <w^2>; 3D_NavierStokes/src/functions.h:128-161,
  2D_NavierStokes/src_RK4/src_gmres/out.F90:2-56);
- the shell spectrum's binning j = INT(|k| + 0.5) (spec.F90:2-57) on single
  modes whose |k| sits just below, just above and exactly on a half-integer;
- the DNS xi-forcing (3D_NavierStokes/src/kernels.h:342-343) and the GMRES
  eta/zeta-forcing (src_gmres/kernels.h:234-238) by their closed-form laminar
  fixed points w = F / (nu |k|^2): both forced pairs sit on one shell
  |k| = KF sqrt 2, the xi case is a 2D flow in (y, z) whose vorticity is a
  Laplacian eigenfunction and the eta/zeta case a unidirectional shear flow
  in x, so the nonlinear term vanishes and Heun + CN keeps the state;
- the configuration IC law E(k) ~ k^4 exp(-2 (k/4)^2) by a fit of the binned
  spectrum (binned here in numpy) over several seeds.
"""
import numpy as np
import pytest

import oracle_lib as O


def full_spectrum(g, f):
    """rfftn-layout array of a stored half-space field, with the kx = 0 plane's
    Hermitian partners filled in (the reference's conj, kernels.h:113-163)."""
    kx, ky, kz, L = O.tables(g)
    ikx = np.rint(kx / g.alpha).astype(int)
    iky = ky.astype(int)
    m = L == 1
    h = m & (ikx == 0)
    if g.dim == 2:
        F = np.zeros((g.ny, g.nx // 2 + 1), complex)
        F[iky[m] % g.ny, ikx[m]] = f[m]
        F[(-iky[h]) % g.ny, 0] = np.conj(f[h])
        return F
    ikz = kz.astype(int)
    F = np.zeros((g.nz, g.ny, g.nx // 2 + 1), complex)
    F[ikz[m] % g.nz, iky[m] % g.ny, ikx[m]] = f[m]
    F[(-ikz[h]) % g.nz, (-iky[h]) % g.ny, 0] = np.conj(f[h])
    return F


def physical(g, f):
    """unnormalised inverse transform (cufftExecZ2D convention) with numpy"""
    shape = (g.ny, g.nx) if g.dim == 2 else (g.nz, g.ny, g.nx)
    return np.fft.irfftn(full_spectrum(g, f), s=shape, axes=tuple(range(len(shape)))) * np.prod(shape)


def test_diag_3d_parseval():
    g = O.grid3(32, 32, 32)
    st = O.ic_solenoidal(g, 4040, 0.7)
    st = O.NS3D(32, v2=0.01, dt=0.002, force=3).run(st, 10)      # a developed state
    kx, ky, kz, L = O.tables(g)
    k2 = np.where(L == 1, kx ** 2 + ky ** 2 + kz ** 2, 1.0)
    xi, eta, zet = st
    u = 1j * (ky * zet - kz * eta) / k2
    v = 1j * (kz * xi - kx * zet) / k2
    w = 1j * (kx * eta - ky * xi) / k2
    up = [physical(g, a) for a in (u, v, w)]
    wp = [physical(g, a) for a in (xi, eta, zet)]
    e_phys = 0.5 * np.mean(sum(a ** 2 for a in up))
    z_phys = 0.5 * np.mean(sum(a ** 2 for a in wp))
    e, z, spec = O.NS3D(32).diag(st)
    assert abs(e - e_phys) <= 1e-12 * e_phys, (e, e_phys)
    assert abs(z - z_phys) <= 1e-12 * z_phys, (z, z_phys)
    assert abs(spec.sum() - e) <= 1e-12 * e        # every active shell lies inside the bins
    # the oracle's own inverse transform gives the same fields
    ph = O.NS3D(32).physical(st)
    assert np.abs(ph[3] - up[0]).max() <= 1e-12 * np.abs(up[0]).max()


def test_diag_sums_exactly_rounded_at_128():
    # the diagnostics add ~1.5e5 active modes per field; compensated sums stay
    # within a few ulps of the exactly rounded total (math.fsum), where a plain
    # running sum drifts by ~n*eps
    import math
    g = O.grid3(128, 128, 128)
    st = O.ic_solenoidal(g, 4141, 0.5)
    kx, ky, kz, L = O.tables(g)
    act = L == 1
    kk = kx ** 2 + ky ** 2 + kz ** 2
    wk = np.where(kk > 0.001, kk, 0.001)
    vz = (np.abs(st) ** 2).sum(axis=0)
    e_ex = math.fsum((vz / wk)[act])
    z_ex = math.fsum(vz[act])
    e, z, spec = O.NS3D(128).diag(st)
    assert abs(e - e_ex) <= 4e-16 * e_ex, (e, e_ex)
    assert abs(z - z_ex) <= 4e-16 * z_ex, (z, z_ex)
    j = (np.sqrt(kk) + 0.5).astype(int)
    for b in range(len(spec)):
        sel = act & (j == b)
        want = math.fsum((vz / wk)[sel])
        assert abs(spec[b] - want) <= 4e-16 * max(want, 1e-300), b


def test_diag_2d_parseval():
    g = O.grid2(64, 64)
    z = O.NS2D(64).run(O.ic_ns2d_cfg(g, 77, 1.0), 20)
    kx, ky, _, L = O.tables(g)
    k2 = np.where(L == 1, kx ** 2 + ky ** 2, 1.0)
    psi = z / k2
    u, v = 1j * ky * psi, -1j * kx * psi
    e_phys = 0.5 * np.mean(physical(g, u) ** 2 + physical(g, v) ** 2)
    v_phys = np.mean(physical(g, z) ** 2)
    e, ens, spec = O.NS2D(64).diag(z)
    assert abs(e - e_phys) <= 1e-12 * e_phys
    assert abs(ens - v_phys) <= 1e-12 * v_phys
    assert abs(spec.sum() - e) <= 1e-12 * e


def put(g, field, k, value, st):
    kx, ky, kz, L = O.tables(g)
    i = np.where((np.abs(kx - k[0]) < 1e-12) & (ky == k[1]) & (kz == k[2]) & (L == 1))[0]
    assert len(i) == 1, k
    st[field, i[0]] = value
    return i[0]


@pytest.mark.parametrize("alpha,k,shell", [
    (1.0, (1, 2, 5), 5),          # |k| = sqrt 30 = 5.48 -> 5
    (1.0, (3, 4, 3), 6),          # |k| = sqrt 34 = 5.83 -> 6
    (1.0, (0, 4, 0), 4),          # on a shell
    (0.5, (0.5, 1, 1), 2),        # |k| = 1.5 exactly: INT(2.0) = 2
    (0.5, (1.5, 0, 4), 4),        # |k| = sqrt 18.25 = 4.27 -> 4
])
def test_shell_spectrum_bins_single_modes(alpha, k, shell):
    ny = 32
    g = O.grid3(int(ny / alpha), ny, ny, alpha)
    st = np.zeros((3, g.nkt), complex)
    put(g, 0, k, 0.3 + 0.4j, st)
    put(g, 2, k, -0.2j, st)
    orc = O.NS3D(ny, alpha=alpha, nx=int(ny / alpha), ny=ny, nz=ny)
    e, z, spec = orc.diag(st)
    kk = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    want = (0.25 + 0.04) / kk
    assert abs(spec[shell] - want) <= 1e-15 and abs(np.delete(spec, shell)).max() == 0.0
    assert abs(e - want) <= 1e-15 and abs(z - 0.29) <= 1e-15


def test_shell_spectrum_bins_2d():
    g = O.grid2(64, 64)
    z = np.zeros(g.nkt, complex)
    kx, ky, _, L = O.tables(g)
    i = np.where((kx == 3) & (ky == -5) & (L == 1))[0][0]    # |k| = sqrt 34 -> 6
    z[i] = 2.0
    e, v, spec = O.NS2D(64).diag(z)
    assert spec[6] == 4.0 / 34 and np.delete(spec, 6).max() == 0.0 and v == 8.0


@pytest.mark.parametrize("mode", ["dns", "gmres"])
def test_forcing_laminar_fixed_points(mode):
    n, kf, nu = 32, 4.0, 0.05
    g = O.grid3(n, n, n)
    k2 = 2 * kf * kf
    st = np.zeros((3, g.nkt), complex)
    if mode == "dns":
        A = -0.25 * kf * kf                      # xi.re += A at (0,KF,KF), -= A at (0,KF,-KF)
        put(g, 0, (0, kf, kf), A / (nu * k2), st)
        put(g, 0, (0, kf, -kf), -A / (nu * k2), st)
        force = 1
    else:
        A = -0.25 * kf                           # eta.im += A, zet.im -= A at (0,KF,KF);
        put(g, 1, (0, kf, kf), 1j * A / (nu * k2), st)      # eta.im += A, zet.im += A at (0,KF,-KF)
        put(g, 2, (0, kf, kf), -1j * A / (nu * k2), st)
        put(g, 1, (0, kf, -kf), 1j * A / (nu * k2), st)
        put(g, 2, (0, kf, -kf), 1j * A / (nu * k2), st)
        force = 2
    kx, ky, kz, L = O.tables(g)
    assert np.abs(kx * st[0] + ky * st[1] + kz * st[2]).max() < 1e-15    # divergence-free
    orc = O.NS3D(n, v2=nu, dt=0.005, force=force, kf=kf)
    # or3_nonlin returns N(w) + F; at the fixed point N = 0 and F = nu |k|^2 w
    assert np.abs(orc.nonlin(st) - nu * k2 * st).max() <= 1e-13 * nu * k2 * np.abs(st).max()
    after = orc.run(st, 40)
    assert np.abs(after - st).max() <= 1e-12 * np.abs(st).max()
    # the other forcing mode does not hold this state
    other = O.NS3D(n, v2=nu, dt=0.005, force=3 - force, kf=kf).run(st, 40)
    assert np.abs(other - st).max() > 1e-3 * np.abs(st).max()


def test_ic_solenoidal_spectrum_law():
    # E(k) ~ k^4 exp(-2 (k/4)^2) over shells of ~4 pi k^2 modes means a per-mode amplitude
    # |u_hat|^2 ~ k^2 exp(-2 (k/4)^2) (in expectation: the Gaussian's two transverse
    # components).  The per-mode ratio to that law, averaged over each shell and 6 seeds
    # (binned here in numpy), is flat over shells 2..14; a k^2 law or exp(-(k/4)^2)
    # would make it vary 49-fold or 200-fold.
    n = 64
    g = O.grid3(n, n, n)
    kx, ky, kz, L = O.tables(g)
    m = L == 1
    kk2 = np.where(m, kx ** 2 + ky ** 2 + kz ** 2, 1.0)
    j = np.floor(np.sqrt(kk2) + 0.5).astype(int)
    num = np.zeros(j.max() + 1)
    cnt = np.bincount(j[m], minlength=num.size)
    for seed in range(6):
        st = O.ic_solenoidal(g, 1000 + seed, 0.5)
        eu = np.where(m, (np.abs(st) ** 2).sum(0) / kk2, 0.0)          # |u_hat|^2 = |w_hat|^2/|k|^2
        num += np.bincount(j, weights=np.where(m, eu / (kk2 * np.exp(-2 * kk2 / 16.0)), 0.0), minlength=num.size)
    ratio = num[2:15] / cnt[2:15]
    assert ratio.max() / ratio.min() < 1.25, ratio / ratio.mean()
