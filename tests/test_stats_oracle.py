"""Pins of the oracle's run statistics (oracle/or_stats.c) by closed forms on
states of one or two modes: the computeStats sums of both flavours (including
the forcing-mode input and the buoyancy flux phase convention), the running
averages, the RiG plane sums of sinusoidal fields, and the spectra binning
(negative vertical wavenumbers in bin 0, as max(k_v, 1e-3) puts them)."""
import numpy as np

import oracle_lib as O


def one_mode(g, nf, kvec, vals):
    """state with the given complex values (one per field) at mode kvec, else 0"""
    kx, ky, kz, L = O.tables(g)
    i = int(np.nonzero((kx == kvec[0]) & (ky == kvec[1]) & (kz == kvec[2]))[0][0])
    s = np.zeros((nf, g.nkt), np.complex128)
    s[:, i] = vals
    assert L[i] == 1
    return s, i


def test_ns3d_sums_of_one_mode_and_running_average():
    g = O.grid3(32, 32, 32)
    k = (1.0, 2.0, -3.0)
    wk = 1 + 4 + 9
    X, E, Z = 0.3 - 0.2j, 0.5 + 0.1j, -0.4 + 0.7j
    s, _ = one_mode(g, 3, k, [X, E, Z])
    t = np.zeros((3, g.nkt), np.complex128)
    v = O.stats_ns3d(g, s, t, 0, -4.0, 4.0)
    vz = abs(X) ** 2 + abs(E) ** 2 + abs(Z) ** 2
    assert np.isclose(v[0], vz) and np.isclose(v[1], vz / wk)
    # first update: the average is the state itself
    assert np.isclose(v[2], vz / wk) and np.isclose(v[3], vz)
    assert v[4] == 0.0
    assert np.allclose(v[5:8], [abs(X) ** 2, abs(E) ** 2, abs(Z) ** 2])
    # spectral velocity of the reference's formulas (functions.h:158-160)
    kx, ky, kz = k
    U = ((ky * Z.imag - kz * E.imag) / wk) ** 2 + ((kz * E.real - ky * Z.real) / wk) ** 2
    assert np.isclose(v[8], U)
    # second update with count 1: <z> = (z0 + z1)/2
    s2 = s * (1 - 2j)
    v2 = O.stats_ns3d(g, s2, t, 1, -4.0, 4.0)
    avg = (s + s2) / 2
    assert np.allclose(t, avg)
    assert np.isclose(v2[3], np.sum(np.abs(avg) ** 2))


def test_ns3d_input_at_the_forcing_modes():
    g = O.grid3(32, 32, 32)
    kf, amp = 4.0, -4.0
    s1, i1 = one_mode(g, 3, (0.0, kf, kf), [0.7 + 0.2j, 0, 0])
    s2, i2 = one_mode(g, 3, (0.0, kf, -kf), [0.25 - 0.5j, 0, 0])
    t = np.zeros((3, g.nkt), np.complex128)
    v = O.stats_ns3d(g, s1 + s2, t, 0, amp, kf)
    assert np.isclose(v[4], amp * 0.7 - amp * 0.25)


def test_bous_buoyancy_flux_phase_and_density_sums():
    g = O.grid3(32, 32, 32)
    # mode k = (0,0,1), vorticity (X, 0, 0): the reference's v = (X.im, -X.re) = -i X,
    # w = 0; with theta = pi/2 the flux is ri |v| |rho| cos(arg rho - arg v) = ri Re(rho conj(v))
    ri, th = 0.7, np.pi / 2
    for rho, want in ((1j, -ri), (1.0, 0.0), (-1j, ri), (2.0 + 2.0j, -2.0 * ri)):
        s, _ = one_mode(g, 4, (0.0, 0.0, 1.0), [1.0, 0, 0, rho])
        t = np.zeros((4, g.nkt), np.complex128)
        v = O.stats_bous(g, s, t, 0, -2.0, 4.0, ri, th)
        assert abs(v[11] - want) < 1e-12, (rho, v[11])
        assert np.isclose(v[5], abs(rho) ** 2) and np.isclose(v[6], abs(rho) ** 2 * 1.0)
        assert v[7] == 0.0 and v[8] == 0.0
    # ped1, ped2 weights and the zeta-mode input -ampfor Re zeta / kf
    s, _ = one_mode(g, 4, (2.0, 3.0, 0.0), [0, 0, 0, 0.5])
    s3, _ = one_mode(g, 4, (0.0, 4.0, 0.0), [0, 0, 0.8 - 0.1j, 0])
    t = np.zeros((4, g.nkt), np.complex128)
    v = O.stats_bous(g, s + s3, t, 0, -2.0, 4.0, ri, th)
    assert np.isclose(v[7], 4 * 0.25) and np.isclose(v[8], 9 * 0.25)
    assert np.isclose(v[4], 2.0 * 0.8 / 4.0)


def test_rig_of_sinusoids():
    nx = ny = nz = 16
    z = 2 * np.pi * np.arange(nz) / nz
    A, B, C = 0.3, 0.8, -0.5
    shape = (nz, ny, nx)
    RO = np.broadcast_to((A * np.sin(z))[:, None, None], shape).copy()
    U = np.broadcast_to((B * np.sin(z))[:, None, None], shape).copy()
    V = np.broadcast_to((C * np.cos(z))[:, None, None], shape).copy()
    W = np.zeros(shape)
    na, sa = O.rig(nx, ny, nz, np.pi / 2, U, V, W, RO)
    h = 2 * np.pi / nz
    f = nz * np.sin(h) / (2 * np.pi)          # central difference of sin: cos(z) f
    assert np.allclose(na, nx * ny * (A * f * np.cos(z) + 1.0))
    assert np.allclose(sa, nx * ny * ((B * f * np.cos(z)) ** 2 + (C * f * np.sin(z)) ** 2))
    # theta < pi/4: planes in y, interior planes by the same closed form in y
    y = 2 * np.pi * np.arange(ny) / ny
    RO = np.broadcast_to((A * np.sin(y))[None, :, None], shape).copy()
    U = np.broadcast_to((B * np.sin(y))[None, :, None], shape).copy()
    na, sa = O.rig(nx, ny, nz, 0.1, U, W, W, RO)
    inner = slice(1, ny - 1)
    assert np.allclose(na[inner], (nx * nz * (A * f * np.cos(y) + 1.0))[inner])
    assert np.allclose(sa[inner], (nx * nz * (B * f * np.cos(y)) ** 2)[inner])
    # the last plane's upper neighbour is plane 1 (out_low.h:31)
    want = nx * nz * (0.5 * ny * (A * np.sin(y[1]) - A * np.sin(y[ny - 2])) / (2 * np.pi) + 1.0)
    assert np.isclose(na[ny - 1], want)


def test_spectra_bins_of_one_mode():
    g = O.grid3(32, 32, 32)
    s, _ = one_mode(g, 3, (3.0, 4.0, -2.0), [1.0, 0, 0])
    h, v, size = O.spectra(g, s, np.pi / 2)
    wk = 9 + 16 + 4
    assert np.isclose(h[5], 1.0 / wk) and np.isclose(v[0], 1.0 / wk)   # |k_h| = 5, k_v = max(-2, 1e-3) -> 0
    # the bin count runs over every stored mode, |k_h| up to sqrt(KTX^2 + KTY^2)
    assert np.count_nonzero(h) == 1 and np.count_nonzero(v) == 1 and size == int(np.sqrt(2 * 10 ** 2) + 0.5)
    # theta < pi/4: horizontal (kx, kz), vertical ky
    h, v, size = O.spectra(g, s, 0.2)
    assert np.isclose(h[int(np.sqrt(13) + 0.5)], 1.0 / wk) and np.isclose(v[4], 1.0 / wk)
