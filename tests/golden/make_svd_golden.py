"""Writes tests/golden/svd_golden.npz: random upper-Hessenberg (k+1) x k
matrices (the GMRES shapes) and their singular values from LAPACK dgesdd via
scipy.linalg.svd(lapack_driver='gesdd').  Run once in this container:
    python tests/golden/make_svd_golden.py"""
import os

import numpy as np
import scipy.linalg as sl

rng = np.random.default_rng(1234)
out = {}
for k in (1, 2, 5, 12, 25, 40):
    H = np.triu(rng.standard_normal((k + 1, k)), -1)
    H[np.arange(1, k + 1), np.arange(k)] = np.abs(H[np.arange(1, k + 1), np.arange(k)]) + 1e-3
    U, s, Vt = sl.svd(H, lapack_driver="gesdd")
    out["H_%d" % k] = H
    out["s_%d" % k] = s
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "svd_golden.npz"), **out)
