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
