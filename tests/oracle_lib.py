"""ctypes wrapper around the FP64 CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The oracle restates the reference scheme
(see oracle/oracle.h for the file:line map) and is the checker for the HIP
product path, never part of it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8"])
        _lib = C.CDLL(LIB)
        _declare(_lib)
    return _lib


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("alpha", C.c_double), ("ktx", C.c_int), ("kty", C.c_int), ("ktz", C.c_int),
                ("iktx", C.c_int), ("ikty", C.c_int), ("iktz", C.c_int), ("nx2", C.c_int),
                ("nkt", C.c_long), ("nr", C.c_long), ("n2", C.c_long)]


P = C.c_void_p
D = C.c_double
I = C.c_int


def _declare(L):
    L.or_fft_c2c.argtypes = [P, I, I]
    L.or_fft_r2c.argtypes = [P, P, I]
    L.or_fft_c2r.argtypes = [P, P, I]
    L.or_fft3_c2r.argtypes = [P, P, I, I, I]
    L.or_fft3_r2c.argtypes = [P, P, I, I, I]
    L.or_grid3.argtypes = [C.POINTER(Grid), I, I, I, D]
    L.or_grid2.argtypes = [C.POINTER(Grid), I, I, D]
    L.or_tables.argtypes = [C.POINTER(Grid), P, P, P, P]
    L.or_count_active.argtypes = [C.POINTER(Grid)]
    L.or_count_active.restype = C.c_long
    L.or_nutab.argtypes = [C.POINTER(Grid), D, D, P, P, P]
    L.or_ikf_ikn.argtypes = [C.POINTER(Grid), P, P]
    L.or_rng_seed.argtypes = [I]
    L.or_ranno0.restype = D
    L.or_pm_next.argtypes = [I]
    L.or_pm_next.restype = I
    L.or_ic_ns3d_ref.argtypes = [C.POINTER(Grid), I, D, P, P, P]
    L.or_ic_ns2d_ref.argtypes = [C.POINTER(Grid), I, D, P]
    L.or_ic_euler_ref.argtypes = [C.POINTER(Grid), I, D, P, P, P]
    L.or_ic_bous_ref.argtypes = [C.POINTER(Grid), I, D, P, P, P, P]
    L.or_ic_solenoidal.argtypes = [C.POINTER(Grid), C.c_uint64, D, D, P, P, P, P]
    L.or_ic_ns2d_cfg.argtypes = [C.POINTER(Grid), C.c_uint64, D, P]
    L.or_gauss.argtypes = [C.c_uint64, C.c_uint64]
    L.or_gauss.restype = D
    L.or3_create.argtypes = [I, I, I, D, I, D, D, I, D, D]
    L.or3_create.restype = P
    L.or3_destroy.argtypes = [P]
    L.or3_run.argtypes = [P, P, P, P, I]
    L.or3_nonlin.argtypes = [P, P, P, P, P, P, P]
    L.or3_physical.argtypes = [P, P, P, P, P]
    L.or3_diag.argtypes = [P, P, P, P, P, P, P]
    L.or2_create.argtypes = [I, I, D, D, D, I, D]
    L.or2_create.restype = P
    L.or2_destroy.argtypes = [P]
    L.or2_run.argtypes = [P, P, I]
    L.or2_nonlin.argtypes = [P, P, P]
    L.or2_diag.argtypes = [P, P, P, P, P]
    L.or_set_threads.argtypes = [I]
    L.or_get_threads.restype = I
    L.or_stats_ns3d.argtypes = [C.c_long] + [P] * 10 + [D, D, D, P]
    L.or_stats_bous.argtypes = [C.c_long] + [P] * 12 + [D, D, D, D, D, P]
    L.or_rig.argtypes = [I, I, I, D, P, P, P, P, P, P]
    L.or_spectra.argtypes = [C.c_long, I, P, P, P, P, P, P, D, P, P]
    L.or_spectra.restype = I
    L.or_rec_new.argtypes = [I, I, I, I, I, I, I, D, P, I, I, I, I, I, I, D, D, D, D, D, P]
    L.or_rec_new.restype = P
    L.or_rec_free.argtypes = [P]
    L.or_rec_update.argtypes = [P, C.c_long, D, P, P, P, P]
    L.or_rec_update.restype = I
    L.or_rec_last.argtypes = [P, P, P, I]
    L.or_rec_last.restype = I
    if hasattr(L, "orb_create"):
        L.orb_create.argtypes = [I, I, I, D, D, D, D, D, D, D, D]
        L.orb_create.restype = P
        L.orb_destroy.argtypes = [P]
        L.orb_run.argtypes = [P, P, P, P, P, I]
        L.orb_rhs.argtypes = [P] + [P] * 8


def ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


# ---------------- thin helpers ----------------
def fft_c2c(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().or_fft_c2c(ptr(x), x.size, sign)
    return x


def fft_r2c(r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(r.size // 2 + 1, np.complex128)
    lib().or_fft_r2c(ptr(r), ptr(out), r.size)
    return out


def fft_c2r(X, n):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    r = np.zeros(n, np.float64)
    lib().or_fft_c2r(ptr(X), ptr(r), n)
    return r


def fft3_c2r(F, nx, ny, nz):
    F = np.ascontiguousarray(F, dtype=np.complex128).copy()
    r = np.zeros((nz, ny, nx), np.float64)
    lib().or_fft3_c2r(ptr(F), ptr(r), nx, ny, nz)
    return r


def fft3_r2c(r):
    nz, ny, nx = r.shape
    r = np.ascontiguousarray(r, dtype=np.float64)
    F = np.zeros((nz, ny, nx // 2 + 1), np.complex128)
    lib().or_fft3_r2c(ptr(r), ptr(F), nx, ny, nz)
    return F


def grid3(nx, ny, nz, alpha=1.0):
    g = Grid()
    lib().or_grid3(C.byref(g), nx, ny, nz, alpha)
    return g


def grid2(nx, ny, alpha=1.0):
    g = Grid()
    lib().or_grid2(C.byref(g), nx, ny, alpha)
    return g


def tables(g):
    n = g.nkt
    kx, ky, kz = (np.zeros(n) for _ in range(3))
    L = np.zeros(n, np.int32)
    lib().or_tables(C.byref(g), ptr(kx), ptr(ky), ptr(kz), ptr(L))
    return kx, ky, kz, L


def nutab(g, v2, dt):
    n = g.nkt
    nu, nuzn, nu1 = (np.zeros(n) for _ in range(3))
    lib().or_nutab(C.byref(g), v2, dt, ptr(nu), ptr(nuzn), ptr(nu1))
    return nu, nuzn, nu1


def ikf_ikn(g):
    ikf = np.zeros(g.n2, np.int32)
    ikn = np.zeros(g.nkt, np.int32)
    lib().or_ikf_ikn(C.byref(g), ptr(ikf), ptr(ikn))
    return ikf, ikn


def ic_ns3d_ref(g, iseed, ampv=1.0):
    s = np.zeros((3, g.nkt), np.complex128)
    lib().or_ic_ns3d_ref(C.byref(g), iseed, ampv, ptr(s[0]), ptr(s[1]), ptr(s[2]))
    return s


def ic_euler_ref(g, iseed, ampv=1.0):
    s = np.zeros((3, g.nkt), np.complex128)
    lib().or_ic_euler_ref(C.byref(g), iseed, ampv, ptr(s[0]), ptr(s[1]), ptr(s[2]))
    return s


def ic_bous_ref(g, iseed, ampv=1.0):
    s = np.zeros((4, g.nkt), np.complex128)
    lib().or_ic_bous_ref(C.byref(g), iseed, ampv, ptr(s[0]), ptr(s[1]), ptr(s[2]), ptr(s[3]))
    return s


def ic_ns2d_ref(g, iseed, ampv=1.0):
    z = np.zeros(g.nkt, np.complex128)
    lib().or_ic_ns2d_ref(C.byref(g), iseed, ampv, ptr(z))
    return z


def ic_solenoidal(g, seed, energy=0.5, rho_scale=None):
    s = np.zeros((4 if rho_scale is not None else 3, g.nkt), np.complex128)
    rho = ptr(s[3]) if rho_scale is not None else None
    lib().or_ic_solenoidal(C.byref(g), seed, energy, rho_scale or 0.0,
                           ptr(s[0]), ptr(s[1]), ptr(s[2]), rho)
    return s


def ic_ns2d_cfg(g, seed, ampv=1.0):
    z = np.zeros(g.nkt, np.complex128)
    lib().or_ic_ns2d_cfg(C.byref(g), seed, ampv, ptr(z))
    return z


class NS3D:
    """Oracle 3D stepper. flavour 0 = Navier-Stokes Heun+CN, 1 = Euler."""

    def __init__(self, n, v2=0.01, dt=0.002, flavour=0, force=3, kf=4.0, ampfor=None,
                 alpha=1.0, nx=None, ny=None, nz=None):
        nx = nx or int(n / alpha)
        ny = ny or n
        nz = nz or n
        if ampfor is None:
            ampfor = {0: 0.0, 1: -0.25 * kf * kf, 2: -0.25 * kf, 3: -0.5 * kf}[force]
        self.h = lib().or3_create(nx, ny, nz, alpha, flavour, v2, dt, force, kf, ampfor)
        self.g = grid3(nx, ny, nz, alpha)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or3_destroy(self.h)
            self.h = None

    def run(self, state, nsteps):
        s = np.ascontiguousarray(state, dtype=np.complex128).copy()
        lib().or3_run(self.h, ptr(s[0]), ptr(s[1]), ptr(s[2]), nsteps)
        return s

    def nonlin(self, state):
        s = np.ascontiguousarray(state, dtype=np.complex128)
        out = np.zeros_like(s)
        lib().or3_nonlin(self.h, ptr(s[0]), ptr(s[1]), ptr(s[2]), ptr(out[0]), ptr(out[1]), ptr(out[2]))
        return out

    def physical(self, state):
        s = np.ascontiguousarray(state, dtype=np.complex128)
        out = np.zeros((6, self.g.nz, self.g.ny, self.g.nx))
        lib().or3_physical(self.h, ptr(s[0]), ptr(s[1]), ptr(s[2]), ptr(out))
        return out

    def diag(self, state):
        s = np.ascontiguousarray(state, dtype=np.complex128)
        e, z = C.c_double(), C.c_double()
        spec = np.zeros(self.g.kty + 2)
        lib().or3_diag(self.h, ptr(s[0]), ptr(s[1]), ptr(s[2]), C.byref(e), C.byref(z), ptr(spec))
        return e.value, z.value, spec


class NS2D:
    def __init__(self, n, v2=1.0 / 40, dt=0.005, alpha=1.0, kf=4, famp=-2.0, nx=None):
        nx = nx or int(n / alpha)
        self.h = lib().or2_create(nx, n, alpha, v2, dt, kf, famp)
        self.g = grid2(nx, n, alpha)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or2_destroy(self.h)
            self.h = None

    def run(self, z, nsteps):
        z = np.ascontiguousarray(z, dtype=np.complex128).copy()
        lib().or2_run(self.h, ptr(z), nsteps)
        return z

    def nonlin(self, z):
        z = np.ascontiguousarray(z, dtype=np.complex128)
        out = np.zeros_like(z)
        lib().or2_nonlin(self.h, ptr(z), ptr(out))
        return out

    def diag(self, z):
        z = np.ascontiguousarray(z, dtype=np.complex128)
        e, v = C.c_double(), C.c_double()
        spec = np.zeros(self.g.kty + 2)
        lib().or2_diag(self.h, ptr(z), C.byref(e), C.byref(v), ptr(spec))
        return e.value, v.value, spec


class BOUS:
    """Oracle 3D Boussinesq RK4/CN stepper (theta in radians, as the reference)."""

    def __init__(self, n, v2=1 / 400.0, sc=1.0, ri=1.0, theta=np.pi / 2, dt=0.002, kf=4.0, ampfor=None,
                 alpha=1.0, nx=None, ny=None, nz=None):
        nx = nx or int(n / alpha)
        ny = ny or n
        nz = nz or n
        if ampfor is None:
            ampfor = -0.5 * kf
        self.h = lib().orb_create(nx, ny, nz, alpha, v2, sc, ri, theta, dt, kf, ampfor)
        self.g = grid3(nx, ny, nz, alpha)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orb_destroy(self.h)
            self.h = None

    def run(self, state, nsteps):
        s = np.ascontiguousarray(state, dtype=np.complex128).copy()
        lib().orb_run(self.h, ptr(s[0]), ptr(s[1]), ptr(s[2]), ptr(s[3]), nsteps)
        return s

    def rhs(self, state):
        s = np.ascontiguousarray(state, dtype=np.complex128)
        out = np.zeros_like(s)
        lib().orb_rhs(self.h, *[ptr(s[i]) for i in range(4)], *[ptr(out[i]) for i in range(4)])
        return out


class Rec:
    """Oracle recurrence search (oracle/or_rec.c): rule 0 3D Navier-Stokes,
    1 Boussinesq, 2 2D.  state arrays are (nf, nkt) complex."""

    def __init__(self, rule, g, L, state0, store_size=500, norm_width=8, nshift=(30, 2, 2), nout=10, dt=0.01,
                 tstart=0.0, threshold=1.0, ri=0.0, v2=0.01):
        s0 = np.ascontiguousarray(state0, dtype=np.complex128)
        self.nf = s0.shape[0] if s0.ndim == 2 else 1
        self.nkt = g.nkt
        self.ns = int(np.prod(nshift))
        self.store_size = store_size
        Lb = np.ascontiguousarray(np.asarray(L).astype(np.int8))
        two = g.dim == 2
        self.h = lib().or_rec_new(rule, self.nf, g.iktx, g.ikty, 1 if two else g.iktz, g.kty, 0 if two else g.ktz,
                                  g.alpha, ptr(Lb), store_size, norm_width, nshift[0], nshift[1], nshift[2], nout,
                                  dt, tstart, threshold, ri, v2, ptr(s0))
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_rec_free(self.h)
            self.h = None

    def update(self, nt, dt, state, want_state=False):
        """Returns (guess or None, guess_state or None, check): guess is a dict of the
        record fields, check of the best match of this check."""
        s = np.ascontiguousarray(state, dtype=np.complex128)
        g = np.zeros(10)
        c = np.zeros(9)
        gs = np.zeros((self.nf, self.nkt), np.complex128) if want_state else None
        r = lib().or_rec_update(self.h, nt, dt, ptr(s), ptr(g), ptr(gs) if want_state else None, ptr(c))
        chk = dict(checked=int(c[0]), compared=int(c[1]), store=int(c[2]), count=int(c[3]), residual=c[4],
                   sx=c[5], sy=c[6], sz=c[7], time=c[8])
        if r != 1:
            return None, None, chk
        guess = dict(index=int(g[0]), store=int(g[1]), time=g[2], v2=g[3], period=g[4], sx=g[5], sy=g[6], sz=g[7],
                     residual=g[8], time_under=g[9])
        return guess, gs, chk

    def last(self):
        res = np.zeros(self.store_size * self.ns)
        slots = np.zeros(self.store_size, np.int32)
        nj = lib().or_rec_last(self.h, ptr(res), ptr(slots), self.store_size)
        return res[:nj * self.ns].reshape(nj, self.ns), slots[:nj]


def stats_ns3d(g, state, tavg, count, ampfor, kf):
    """raw computeStats sums (3D Navier-Stokes); tavg (3, nkt) updated in place"""
    kx, ky, kz, L = tables(g)
    st = np.ascontiguousarray(state, dtype=np.complex128)
    out = np.zeros(11)
    lib().or_stats_ns3d(g.nkt, ptr(kx), ptr(ky), ptr(kz), ptr(L), ptr(st[0]), ptr(st[1]), ptr(st[2]),
                        ptr(tavg[0]), ptr(tavg[1]), ptr(tavg[2]), float(count), ampfor, kf, ptr(out))
    return out


def stats_bous(g, state, tavg, count, ampfor, kf, ri, theta):
    kx, ky, kz, L = tables(g)
    st = np.ascontiguousarray(state, dtype=np.complex128)
    out = np.zeros(12)
    lib().or_stats_bous(g.nkt, ptr(kx), ptr(ky), ptr(kz), ptr(L), ptr(st[0]), ptr(st[1]), ptr(st[2]), ptr(st[3]),
                        ptr(tavg[0]), ptr(tavg[1]), ptr(tavg[2]), ptr(tavg[3]), float(count), ampfor, kf, ri, theta,
                        ptr(out))
    return out


def rig(nx, ny, nz, theta, U, V, W, RO):
    npl = ny if theta < 2 * np.pi / 8 else nz
    na, sa = np.zeros(npl), np.zeros(npl)
    f = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (U, V, W, RO)]
    lib().or_rig(nx, ny, nz, theta, ptr(f[0]), ptr(f[1]), ptr(f[2]), ptr(f[3]), ptr(na), ptr(sa))
    return na, sa


def spectra(g, tavg, theta):
    kx, ky, kz, L = tables(g)
    t = np.ascontiguousarray(tavg, dtype=np.complex128)
    h, v = np.zeros(g.nx + 1), np.zeros(g.nx + 1)
    size = lib().or_spectra(g.nkt, g.nx, ptr(kx), ptr(ky), ptr(kz), ptr(t[0]), ptr(t[1]), ptr(t[2]), theta, ptr(h),
                            ptr(v))
    return h, v, size
