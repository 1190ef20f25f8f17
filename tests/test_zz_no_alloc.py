"""psg_diag_noalloc (the diagnostics entry point that allocates nothing)
against psg_diag, through the C-ABI, on two calls in a row: a 64^3
Navier-Stokes state, then a smaller 32^3 one whose fewer partial rows must not
pick up the first call's leftovers in the shared device array; and a 2D
state (enstrophy doubled on the device)."""
import ctypes as C

import numpy as np
import pytest

import psgpu_amd as P

pytestmark = pytest.mark.gpu


def both(s, member=0):
    L = P.lib()
    e0, z0 = C.c_double(), C.c_double()
    e1, z1 = C.c_double(), C.c_double()
    kty = (s.ikty - 1) // 2
    sp0, sp1 = np.zeros(kty + 2), np.zeros(kty + 2)
    assert L.psg_diag(s.h, member, C.byref(e0), C.byref(z0), sp0.ctypes.data_as(C.c_void_p)) == 0
    rc = L.psg_diag_noalloc(s.h, member, C.byref(e1), C.byref(z1), sp1.ctypes.data_as(C.c_void_p))
    assert rc == 0, L.psg_last_error()
    return (e0.value, z0.value, sp0), (e1.value, z1.value, sp1)


def check(a, b):
    (e0, z0, s0), (e1, z1, s1) = a, b
    assert abs(e1 - e0) <= 1e-15 * abs(e0) and abs(z1 - z0) <= 1e-15 * abs(z0), (e0, e1, z0, z1)
    assert np.abs(s1 - s0).max() <= 1e-15 * np.abs(s0).max()
    print("bitwise equal:", e0 == e1 and z0 == z1 and np.array_equal(s0, s1), flush=True)


def test_diag_noalloc_matches_diag_twice_smaller_second(gpu):
    for n, seed in ((64, 901), (32, 902)):
        g = P.hgrid("ns3d", n, n, n)
        with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, force="kolmogorov") as s:
            s.set_state(P.ic_solenoidal(g, seed, 0.5))
            s.run(3)
            a, b = both(s)
            check(a, b)


def test_diag_noalloc_2d_and_members(gpu):
    g = P.hgrid("ns2d", 128, 128)
    with P.Stepper(flavour="ns2d", n=128, v2=1 / 40, dt=0.005, force="kolmogorov", kf=4.0, ampfor=-2.0,
                   batch=2) as s:
        for m in range(2):
            s.set_state(P.ic_ns2d_cfg(g, 77 + m, 1.0)[None], m)
        s.run(2)
        for m in (1, 0):
            a, b = both(s, m)
            check(a, b)
