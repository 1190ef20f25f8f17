"""CPU pins of the Newton solver's dense algebra: the Hessenberg SVD against
LAPACK dgesdd goldens and orthogonality/reconstruction; the hookstep
trust-region search against the Levenberg-Marquardt closed form."""
import ctypes as C
import os

import numpy as np
import pytest

import psgpu_amd as P

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "svd_golden.npz"))


def svd(A):
    m, n = A.shape
    a = np.asfortranarray(A, dtype=np.float64)
    U, s, V = np.zeros((n, m)), np.zeros(n), np.zeros((n, n))
    assert P.lib().psg_svd(m, n, a.ctypes.data_as(C.c_void_p), U.ctypes.data_as(C.c_void_p),
                           s.ctypes.data_as(C.c_void_p), V.ctypes.data_as(C.c_void_p)) == 0
    return U.T, s, V.T   # column-major storage -> U (m x n), V (n x n)


@pytest.mark.parametrize("k", [1, 2, 5, 12, 25, 40])
def test_svd_matches_gesdd(k):
    H = GOLD["H_%d" % k]
    U, s, V = svd(H)
    assert np.abs(s - GOLD["s_%d" % k]).max() < 1e-13 * GOLD["s_%d" % k].max()
    assert np.abs(U @ np.diag(s) @ V.T - H).max() < 1e-13 * np.abs(H).max()
    assert np.abs(U.T @ U - np.eye(k)).max() < 1e-13
    assert np.abs(V.T @ V - np.eye(k)).max() < 1e-13
    assert (np.diff(s) <= 0).all()


def test_hookstep_is_levenberg_marquardt():
    rng = np.random.default_rng(5)
    k = 12
    H = np.triu(rng.standard_normal((k + 1, k)), -1)
    U, s, V = svd(H)
    b = np.zeros(k + 1); b[0] = 2.0
    p = 2.0 * U[0, :]
    for delta in (10.0, 0.5, 0.05):
        x = np.zeros(k)
        mu = P.lib().psg_hook_mu(k, s.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p), delta,
                                 x.ctypes.data_as(C.c_void_p))
        y = V @ x
        nx = np.linalg.norm(x)
        if mu == 0.0:
            # unconstrained least-squares step already inside the region
            assert nx <= np.sqrt(2) * delta
            assert np.abs(H.T @ (H @ y - b)).max() < 1e-10
        else:
            assert delta / np.sqrt(2) < nx < np.sqrt(2) * delta
            # (H^T H + mu I) y = H^T b : the Levenberg-Marquardt / trust-region step
            assert np.abs((H.T @ H + mu * np.eye(k)) @ y - H.T @ b).max() < 1e-10


def hess_eig(H):
    n = H.shape[0]
    a = np.asfortranarray(H, dtype=np.float64)
    wr, wi = np.zeros(n), np.zeros(n)
    assert P.lib().psg_hess_eig(n, a.ctypes.data_as(C.c_void_p), wr.ctypes.data_as(C.c_void_p),
                                wi.ctypes.data_as(C.c_void_p)) == 0
    return wr + 1j * wi


def match(got, want, tol):
    """every eigenvalue of want paired with a distinct one of got"""
    got = list(got)
    for w in want:
        i = int(np.argmin([abs(g - w) for g in got]))
        assert abs(got[i] - w) <= tol, (w, got[i])
        got.pop(i)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 17, 40, 97])
def test_hess_eig_vs_numpy_random(n):
    # Hessenberg QR eigenvalues against LAPACK's (numpy.linalg.eigvals) on
    # random Hessenberg matrices; conjugate pairs adjacent, positive part first
    rng = np.random.default_rng(100 + n)
    H = np.triu(rng.standard_normal((n, n)), -1)
    ev = hess_eig(H)
    scale = max(1.0, np.abs(H).max())
    match(ev, np.linalg.eigvals(H), 1e-10 * scale * n)
    for i in range(n - 1):
        if ev[i].imag > 0:
            assert ev[i + 1] == np.conj(ev[i])


def test_hess_eig_structured():
    # a Floquet-like spectrum: a few leading multipliers over a decaying bulk,
    # similarity-transformed to Hessenberg form by numpy's QR of a random basis
    rng = np.random.default_rng(7)
    n = 60
    lam = np.concatenate([[1.7, -0.9, 0.5 + 0.6j, 0.5 - 0.6j], 0.05 * rng.random(n - 4)])
    D = np.zeros((n, n))
    i = 0
    while i < n:
        if abs(lam[i].imag) > 0:
            a, b = lam[i].real, lam[i].imag
            D[i:i + 2, i:i + 2] = [[a, b], [-b, a]]
            i += 2
        else:
            D[i, i] = lam[i].real
            i += 1
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    import scipy.linalg as sl
    Hh = sl.hessenberg(Q @ D @ Q.T)
    ev = hess_eig(Hh)
    match(ev, lam, 1e-10)
    # zero subdiagonal (already split) and a Jordan-like defective 2 x 2 block
    H2 = np.diag([3.0, 2.0, 1.0]) + np.diag([0.0, 0.0], -1) + np.triu(np.ones((3, 3)), 1)
    match(hess_eig(H2), [3.0, 2.0, 1.0], 1e-14)
    J = np.array([[2.0, 1.0], [0.0, 2.0]])
    match(hess_eig(J), [2.0, 2.0], 1e-7)
