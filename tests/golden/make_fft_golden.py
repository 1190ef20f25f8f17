"""Writes tests/golden/fft_golden.npz: numpy.fft reference vectors (FP64) for
the oracle's and the GPU path's transforms.  Run once in this container:
    python tests/golden/make_fft_golden.py
Inputs are seeded; outputs are numpy's pocketfft results."""
import os

import numpy as np

rng = np.random.default_rng(20240611)
out = {}
for n in (8, 16, 32, 64, 128, 256, 512):
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = rng.standard_normal(n)
    out["c_in_%d" % n] = x
    out["c_fwd_%d" % n] = np.fft.fft(x)
    out["c_inv_%d" % n] = np.fft.ifft(x) * n
    out["r_in_%d" % n] = r
    out["r_fwd_%d" % n] = np.fft.rfft(r)
for shape in ((8, 16, 32), (16, 16, 16), (32, 8, 16)):
    r = rng.standard_normal(shape)
    key = "x".join(map(str, shape))
    out["r3_in_" + key] = r
    out["r3_fwd_" + key] = np.fft.rfftn(r)
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz"), **out)
