"""CPU pins of the FP64 oracle: golden vectors, reference-held constants,
closed forms.  Each check is chosen so a plausible mistake in the oracle
(wrong sign, dropped term, transposed index) fails it."""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256, 512])
def test_fft_golden(n):
    x = GOLD["c_in_%d" % n]
    tol = 1e-13 * np.abs(GOLD["c_fwd_%d" % n]).max()
    assert np.abs(O.fft_c2c(x, -1) - GOLD["c_fwd_%d" % n]).max() < tol
    assert np.abs(O.fft_c2c(x, +1) - GOLD["c_inv_%d" % n]).max() < tol
    r = GOLD["r_in_%d" % n]
    assert np.abs(O.fft_r2c(r) - GOLD["r_fwd_%d" % n]).max() < tol
    # c2r of rfft(r) is n*r (cuFFT Z2D convention)
    assert np.abs(O.fft_c2r(GOLD["r_fwd_%d" % n], n) - n * r).max() < 1e-12 * n


@pytest.mark.parametrize("key", ["8x16x32", "16x16x16", "32x8x16"])
def test_fft3_golden(key):
    r = GOLD["r3_in_" + key]
    F = GOLD["r3_fwd_" + key]
    assert np.abs(O.fft3_r2c(r) - F).max() < 1e-12 * np.abs(F).max()
    nz, ny, nx = r.shape
    assert np.abs(O.fft3_c2r(F, nx, ny, nz) / r.size - r).max() < 1e-13


def test_fft_closed_forms():
    n = 64
    d = np.zeros(n); d[0] = 1.0
    assert np.allclose(O.fft_r2c(d), np.ones(n // 2 + 1))          # delta -> constant
    j = np.arange(n)
    c = np.cos(2 * np.pi * 5 * j / n)
    X = O.fft_r2c(c)
    assert abs(X[5] - n / 2) < 1e-12 and np.abs(np.delete(X, 5)).max() < 1e-12
    x = np.random.default_rng(3).standard_normal(n)
    X = O.fft_r2c(x)                                                 # Parseval
    e = (abs(X[0]) ** 2 + abs(X[-1]) ** 2 + 2 * (np.abs(X[1:-1]) ** 2).sum()) / n
    assert abs(e - (x ** 2).sum()) < 1e-12 * e


def test_reference_held_sizes():
    # 3D_NavierStokes/src_gmres/kernels.h:12  '#define nkt 310675' (128^3)
    assert O.grid3(128, 128, 128).nkt == 310675
    # 2D 256^2: SzV = 22428 = 2 * active (2D_NavierStokes/src_RK4/src_gmres/io_arpack.F90:6)
    assert 2 * O.lib().or_count_active(O.grid2(256, 256)) == 22428
    # Boussinesq 128x64x64, alpha=0.5: SzV = 288872 = 2*4*active (3D_Boussinesq/src/src_gmres/io_arpack.F90:6)
    assert 8 * O.lib().or_count_active(O.grid3(128, 64, 64, 0.5)) == 288872
    # 256^3 Zk.out header 1,86,171,171 (3D_Euler/bin/Zk.out)
    g = O.grid3(256, 256, 256)
    assert (g.iktx, g.ikty, g.iktz) == (86, 171, 171)


def test_mask_bruteforce():
    for (nx, n) in [(12, 12), (24, 24), (24, 12)]:
        g = O.grid3(nx, n, n, n / nx)
        kx, ky, kz, L = O.tables(g)
        kt = n // 3
        for i in range(g.nkt):
            half = kx[i] > 0 or (kx[i] == 0 and (ky[i] > 0 or (ky[i] == 0 and kz[i] > 0)))
            inside = np.sqrt(kx[i] ** 2 + ky[i] ** 2 + kz[i] ** 2) <= kt - 0.5
            assert L[i] == int(half and inside)
        # exactly one of k, -k is active for every k in the sphere (but 0)
        act = {(kx[i], ky[i], kz[i]) for i in range(g.nkt) if L[i]}
        for (a, b, c) in act:
            assert (-a, -b, -c) not in act


def test_ikf_ikn_inverse():
    for g in (O.grid3(32, 32, 32), O.grid3(32, 16, 16, 0.5), O.grid2(32, 32)):
        ikf, ikn = O.ikf_ikn(g)
        assert (ikf[ikn] == np.arange(g.nkt)).all()
        assert (ikf >= 0).sum() == g.nkt


def test_park_miller_check_value():
    # Park & Miller (1988): seed 1 -> x_10000 = 1043618065
    s = 1
    for _ in range(10000):
        s = O.lib().or_pm_next(s)
    assert s == 1043618065


def test_ranno_reseeds_every_draw():
    # ranno(0) draws with a freshly re-seeded ran1 (seed-1, seed-2, ...), so the
    # n-th draw equals the first draw of a fresh sequence seeded with seed+n-1
    O.lib().or_rng_seed(1000)
    a = [O.lib().or_ranno0() for _ in range(3)]
    O.lib().or_rng_seed(1001)
    b = O.lib().or_ranno0()
    assert a[1] == b
    assert all(-np.pi <= v < np.pi for v in a)


def test_nutab_formula():
    g = O.grid3(16, 16, 16)
    nu, nuzn, nu1 = O.nutab(g, 0.01, 0.002)
    kx, ky, kz, L = O.tables(g)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    on = L == 1
    assert np.allclose(nu[on], 0.01 * k2[on], rtol=1e-14)
    assert np.allclose(nu1[on], 0.002 / (1 + 0.001 * nu[on]), rtol=1e-15)
    assert (nu[~on] == 0).all() and (nu1[~on] == 0).all()


def test_ic_solenoidal():
    g = O.grid3(32, 32, 32)
    s = O.ic_solenoidal(g, 7)
    kx, ky, kz, L = O.tables(g)
    k = np.stack([kx, ky, kz])
    assert np.abs((k * s).sum(0)).max() < 1e-14           # div w = 0
    k2 = np.maximum((k ** 2).sum(0), 1e-3)
    e = (np.abs(s) ** 2).sum(0) / k2
    assert abs(e[L == 1].sum() - 0.5) < 1e-14              # sum |u|^2 = 0.5
    assert np.abs(s[:, L == 0]).max() == 0.0


def test_ic_reference_enstrophy():
    g = O.grid3(32, 32, 32)
    s = O.ic_ns3d_ref(g, 837920, 1.0)
    assert abs(2 * (np.abs(s) ** 2).sum() - 1.0) < 1e-13
    g2 = O.grid2(64, 64)
    z = O.ic_ns2d_ref(g2, 837920, 1.0)
    assert abs(2 * (np.abs(z) ** 2).sum() - 1.0) < 1e-13
    kx, ky, _, L = O.tables(g2)
    assert np.abs(z[kx == 0]).max() == 0.0                  # kx=0 modes zeroed


def _phases(seed, n):
    O.lib().or_rng_seed(seed)
    return np.array([O.lib().or_ranno0() for _ in range(n)])


def _loop_order(g):
    # position of each stored mode in the reference's IKX, IKY, IKZ (IKZ innermost) loop
    ikz, iky, ikx = np.meshgrid(np.arange(g.iktz), np.arange(g.ikty), np.arange(g.iktx), indexing="ij")
    return ((ikx * g.ikty + iky) * g.iktz + ikz).ravel()


def test_ic_euler_reference_law():
    g = O.grid3(16, 16, 16)
    ampv = 0.3
    s = O.ic_euler_ref(g, 837920, ampv)
    kx, ky, kz, L = O.tables(g)
    wk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    assert np.abs(s[:2]).max() == 0.0                                       # xii = eta = 0
    tg = (kx == 1) & (kz == 0) & (np.abs(ky) == 1)
    assert np.array_equal(s[2, tg].real, 0.25 * ky[tg]) and np.abs(s[2, tg].imag).max() == 0.0
    shell = (wk > 1.5) & (wk <= 3.5) & (kz == 1)
    assert shell.sum() > 10 and np.abs(np.abs(s[2, shell]) - ampv).max() < 1e-15
    assert np.abs(s[2, ~(shell | tg)]).max() == 0.0
    # one phase per stored mode, in loop order: the shell phases are those draws
    ph = _phases(837920, g.nkt)[_loop_order(g)]
    assert np.abs(s[2, shell] - ampv * np.exp(1j * ph[shell])).max() < 1e-15


def test_ic_boussinesq_reference_law():
    g = O.grid3(16, 16, 16)
    ampv = 2.0
    s = O.ic_bous_ref(g, 4242, ampv)
    kx, ky, kz, L = O.tables(g)
    wk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    assert np.abs(s[3]).max() == 0.0                                        # rho = 0
    assert np.abs(kx * s[0] + ky * s[1] + kz * s[2]).max() < 1e-15          # k . w = 0
    on = (wk > 2.5) & (wk <= 9.5) & (kz != 0) & (L == 1)
    assert np.abs(s[:, ~on]).max() == 0.0
    a = np.abs(s[0, on])
    assert np.abs(a - a[0]).max() < 1e-15 and np.abs(np.abs(s[1, on]) - a[0]).max() < 1e-15
    # three draws per stored mode (the third unused): xii and eta carry draws 3p and 3p+1
    ph = _phases(4242, 3 * g.nkt)
    p = _loop_order(g)
    assert np.abs(s[0, on] - a[0] * np.exp(1j * ph[3 * p[on]])).max() < 1e-14
    assert np.abs(s[1, on] - a[0] * np.exp(1j * ph[3 * p[on] + 1])).max() < 1e-14
    # the enstrophy that sets the amplitude runs over every stored mode, the
    # conjugate duplicates at kx = 0 included, before the mask is applied
    act = (wk > 2.5) & (wk <= 9.5) & (kz != 0)
    x0 = 0.1 * np.exp(1j * ph[3 * p])
    e0 = 0.1 * np.exp(1j * ph[3 * p + 1])
    z0 = np.where(kz != 0, -(kx * x0 + ky * e0) / np.where(kz != 0, kz, 1), 0)
    ens = 2 * np.sum(act * (np.abs(x0) ** 2 + np.abs(e0) ** 2 + np.abs(z0) ** 2))
    assert abs(a[0] - 0.1 * np.sqrt(ampv / ens)) < 1e-15


def test_single_mode_cn_decay():
    # a single Fourier mode has zero nonlinear term: one Heun+CN step is (1-a)/(1+a)
    s = O.NS3D(16, v2=0.05, dt=0.01, force=0)
    kx, ky, kz, L = O.tables(s.g)
    st = np.zeros((3, s.g.nkt), complex)
    i = np.where((kx == 1) & (ky == 2) & (kz == 0))[0][0]
    k = np.array([1.0, 2.0, 0.0])
    u = np.array([2.0, -1.0, 0.0]) * (0.3 + 0.2j)
    st[:, i] = 1j * np.cross(k, u)
    nu = 10 ** (np.log10(0.05) + 2 * np.log10(np.sqrt(5.0)))
    a = 0.5 * nu * 0.01
    out = s.run(st, 5)
    assert np.abs(out[:, i] - st[:, i] * ((1 - a) / (1 + a)) ** 5).max() < 4e-15
    # 2D RK4 + end-of-step CN
    s2 = O.NS2D(16, v2=0.05, dt=0.01, famp=0.0)
    kx, ky, _, L = O.tables(s2.g)
    z = np.zeros(s2.g.nkt, complex)
    i = np.where((kx == 2) & (ky == -1))[0][0]
    z[i] = 0.7 - 0.1j
    a = 0.5 * 0.01 * 0.05 * 5
    assert abs(s2.run(z, 3)[i] - z[i] * ((1 - a) / (1 + a)) ** 3) < 4e-15


def test_kolmogorov_laminar_fixed_point():
    # u = (Re/KF^2) sin(KF y) x with f=(sin KF y,0,0): zeta(0,KF,0) = -Re/(2 KF)
    for re_ in (10.0, 40.0):
        s = O.NS3D(16, v2=1 / re_, dt=0.01, force=3, kf=4.0)
        kx, ky, kz, L = O.tables(s.g)
        st = np.zeros((3, s.g.nkt), complex)
        st[2, np.where((kx == 0) & (ky == 4) & (kz == 0))[0][0]] = -re_ / 8
        assert np.abs(s.run(st, 10) - st).max() < 1e-13 * re_
    s2 = O.NS2D(16, v2=1 / 40.0, dt=0.01)
    kx, ky, _, L = O.tables(s2.g)
    z = np.zeros(s2.g.nkt, complex)
    z[np.where((kx == 0) & (ky == 4))[0][0]] = -40.0 / 8   # subLam: +0.125/v2 removes it
    assert np.abs(s2.run(z, 10) - z).max() < 1e-12


def test_euler_energy_conservation_order():
    # Galerkin-truncated Euler conserves energy in the dt -> 0 limit.  The
    # reference predictor keeps the cube-truncated modes outside the 2/3 sphere
    # (3D_Euler/src/kernels.h:339-345 has no mask on w*), which feeds an O(dt)
    # term into the corrector: the drift over a fixed time is first order.
    g = O.grid3(16, 16, 16)
    st0 = O.ic_solenoidal(g, 11, 0.5)
    drift = []
    for dt in (0.01, 0.005, 0.0025):
        s = O.NS3D(16, dt=dt, flavour=1, force=0)
        out = s.run(st0, int(round(0.2 / dt)))
        drift.append(abs(s.diag(out)[0] - 0.5))
    assert drift[0] < 2e-3
    assert 1.8 < drift[0] / drift[1] < 2.2 and 1.8 < drift[1] / drift[2] < 2.2


def test_heun_order():
    g = O.grid3(16, 16, 16)
    st0 = O.ic_solenoidal(g, 5, 0.5)
    T = 0.1
    ref = O.NS3D(16, v2=0.02, dt=T / 64, force=3).run(st0, 64)
    errs = []
    for nst in (8, 16):
        out = O.NS3D(16, v2=0.02, dt=T / nst, force=3).run(st0, nst)
        errs.append(np.abs(out - ref).max())
    assert 3.0 < errs[0] / errs[1] < 5.0


def test_nonlinear_term_is_curl_of_product():
    # independent check of the oracle RHS: compute curl(u x w) directly with numpy FFTs
    n = 16
    s = O.NS3D(n, force=0)
    g = s.g
    st = O.ic_solenoidal(g, 3, 0.5)
    kx, ky, kz, L = O.tables(g)
    # full spectral arrays for numpy (zero-padded, Hermitian)
    F = np.zeros((3, n, n, n), complex)
    for f in range(3):
        for i in np.where(L == 1)[0]:
            a, b, c = int(kx[i]), int(ky[i]), int(kz[i])
            F[f, c % n, b % n, a % n] += st[f, i]
            F[f, (-c) % n, (-b) % n, (-a) % n] += np.conj(st[f, i])
    K = np.meshgrid(np.fft.fftfreq(n, 1 / n), np.fft.fftfreq(n, 1 / n), np.fft.fftfreq(n, 1 / n), indexing="ij")
    kzz, kyy, kxx = K
    kvec = np.stack([kxx, kyy, kzz])
    k2 = np.maximum((kvec ** 2).sum(0), 1e-3)
    U = 1j * np.cross(kvec, F, axis=0) / k2
    w = np.fft.ifftn(F, axes=(1, 2, 3)).real * n ** 3
    u = np.fft.ifftn(U, axes=(1, 2, 3)).real * n ** 3
    P = np.fft.fftn(np.cross(u, w, axis=0), axes=(1, 2, 3)) / n ** 3
    C = 1j * np.cross(kvec, P, axis=0)
    want = np.zeros_like(st)
    for i in np.where(L == 1)[0]:
        a, b, c = int(kx[i]), int(ky[i]), int(kz[i])
        want[:, i] = C[:, c % n, b % n, a % n]
    got = s.nonlin(st)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_bous_single_mode_decay_and_fixed_point():
    # single Fourier mode, no buoyancy coupling (u along x, k along z, theta=pi/2):
    # RK4 stages see C=0, the step is (1-a)/(1+a), a = dt nu k^2 / 2
    b = O.BOUS(16, v2=0.05, ri=1.0, theta=np.pi / 2, dt=0.01, ampfor=0.0)
    kx, ky, kz, L = O.tables(b.g)
    st = np.zeros((4, b.g.nkt), complex)
    i = np.where((kx == 0) & (ky == 0) & (kz == 2))[0][0]
    u = np.array([0.3 - 0.1j, 0.0, 0.0])
    st[:3, i] = 1j * np.cross([0.0, 0.0, 2.0], u)
    a = 0.5 * 0.01 * 0.05 * 4
    out = b.run(st, 4)
    assert np.abs(out[:, i] - st[:, i] * ((1 - a) / (1 + a)) ** 4).max() < 1e-15
    assert np.abs(out[3]).max() < 1e-17
    # Kolmogorov laminar state is a fixed point with rho = 0
    re_ = 20.0
    b = O.BOUS(16, v2=1 / re_, ri=1.0, theta=np.pi / 2, dt=0.01, kf=4.0)
    st = np.zeros((4, b.g.nkt), complex)
    st[2, np.where((kx == 0) & (ky == 4) & (kz == 0))[0][0]] = -re_ / 8
    assert np.abs(b.run(st, 5) - st).max() < 1e-13 * re_
    # dz/dt vanishes there
    assert np.abs(b.rhs(st)).max() < 1e-13 * re_


def test_bous_projection_and_rk4_order():
    g = O.grid3(16, 16, 16)
    st = np.concatenate([O.ic_solenoidal(g, 9, 0.5, 0.1)], 0)
    T = 0.08
    ref = O.BOUS(16, v2=0.02, dt=T / 64).run(st, 64)
    kx, ky, kz, L = O.tables(g)
    div = np.abs(kx * ref[0] + ky * ref[1] + kz * ref[2]).max()
    assert div < 1e-13 * np.abs(ref[:3]).max()        # Leray projection keeps div w = 0
    errs = []
    for nst in (8, 16):
        errs.append(np.abs(O.BOUS(16, v2=0.02, dt=T / nst).run(st, nst) - ref).max())
    assert errs[0] / errs[1] > 3.0                      # at least second order (RK4 stages, CN viscous part)


def test_nonlinear_terms_conserve_quadratic_invariants():
    # The 2/3-truncated advection term is conservative: 3D Navier-Stokes / Euler
    # keep the energy sum_L |w|^2/|k|^2; 2D keeps energy and enstrophy.  A wrong
    # sign or index in one component of the product, Biot-Savart or curl breaks it.
    g = O.grid3(16, 16, 16)
    kx, ky, kz, L = O.tables(g)
    k2 = np.maximum(kx ** 2 + ky ** 2 + kz ** 2, 1e-3)
    s = O.NS3D(16, force=0)
    for seed in range(3):
        st = O.ic_solenoidal(g, 200 + seed, 0.5)
        n = s.nonlin(st)
        de = np.sum((np.conj(st) * n).real * L / k2)
        scale = np.sum(np.abs(st) * np.abs(n) * L / k2)
        assert abs(de) < 1e-13 * scale, (de, scale)
    g2 = O.grid2(32, 32)
    kx2, ky2, _, L2 = O.tables(g2)
    k22 = np.maximum(kx2 ** 2 + ky2 ** 2, 1e-3)
    s2 = O.NS2D(32, famp=0.0)
    for seed in range(3):
        z = O.ic_ns2d_cfg(g2, 300 + seed)
        n = s2.nonlin(z)
        scale = np.sum(np.abs(z) * np.abs(n) * L2)
        assert abs(np.sum((np.conj(z) * n).real * L2)) < 1e-13 * scale           # enstrophy
        assert abs(np.sum((np.conj(z) * n).real * L2 / k22)) < 1e-13 * scale     # energy


def test_boussinesq_buoyancy_exchange_conserves_total_energy():
    # With diffusion and forcing off, kinetic energy sum|w|^2/|k|^2 and Ri sum|rho|^2
    # exchange through the buoyancy terms only: -dKE/dt / d(sum|rho|^2)/dt = Ri for
    # any state and any gravity angle theta (radians, as the reference passes it).
    g = O.grid3(16, 16, 16)
    kx, ky, kz, L = O.tables(g)
    k2 = np.maximum(kx ** 2 + ky ** 2 + kz ** 2, 1e-3)
    for theta in (np.pi / 2, 0.3, 90.0):
        b = O.BOUS(16, v2=1e-14, sc=1.0, ri=1.7, theta=theta, dt=0.01, ampfor=0.0)
        for seed in range(3):
            st = O.ic_solenoidal(g, 100 + seed, 0.5, 0.3)
            d = b.rhs(st)
            dke = 2 * np.sum((np.conj(st[:3]) * d[:3]).real * L / k2)
            dpe = 2 * np.sum((np.conj(st[3]) * d[3]).real * L)
            assert abs(-dke / dpe - 1.7) < 1e-8, (theta, seed, -dke / dpe)


def test_nonlinear_term_2d_matches_numpy():
    # independent restatement with numpy FFTs: dw/dt = -(d/dx(u w) + d/dy(v w)),
    # u = d psi/dy, v = -d psi/dx, psi = w/|k|^2, forcing -2 at (0, +-4)
    n = 32
    s = O.NS2D(n, famp=-2.0)
    g = s.g
    z = O.ic_ns2d_cfg(g, 17)
    kx, ky, _, L = O.tables(g)
    F = np.zeros((n, n), complex)
    for i in np.where(L == 1)[0]:
        a, b = int(kx[i]), int(ky[i])
        F[b % n, a % n] += z[i]
        F[(-b) % n, (-a) % n] += np.conj(z[i])
    ky_, kx_ = np.meshgrid(np.fft.fftfreq(n, 1 / n), np.fft.fftfreq(n, 1 / n), indexing="ij")
    k2 = np.maximum(kx_ ** 2 + ky_ ** 2, 1e-3)
    psi = F / k2
    u = np.fft.ifft2(1j * ky_ * psi).real * n * n
    v = np.fft.ifft2(-1j * kx_ * psi).real * n * n
    w = np.fft.ifft2(F).real * n * n
    N = -(1j * kx_ * np.fft.fft2(u * w) + 1j * ky_ * np.fft.fft2(v * w)) / (n * n)
    want = np.zeros_like(z)
    for i in np.where(L == 1)[0]:
        a, b = int(kx[i]), int(ky[i])
        want[i] = N[b % n, a % n]
        if a == 0 and abs(b) == 4:
            want[i] += -2.0
    got = s.nonlin(z)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
