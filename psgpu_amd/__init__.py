"""psgpu-mi355x: MI355X-native pseudospectral steppers (Python host mirror).

The compute path is the HIP library psgpu_amd/_lib/libpsg.so (hand-written
gfx950 kernels behind the C-ABI in include/psg.h).  This module is a thin
ctypes mirror of that C-ABI for tests, benchmarks and Python drivers: there
is no Python or CPU fallback -- if the library is missing, or no GPU is
visible, construction raises.

Reference mapping: a ``Stepper`` replaces one reference ``timestep_cuda_``
call sequence (3D_NavierStokes/src/cuda.cu:18-151 and siblings) with a
persistent context; ``run(n)`` advances every member by n steps.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "_lib")
LIBPATH = os.path.join(LIBDIR, "libpsg.so")

NS2D, NS3D, EULER3D, BOUS3D = 1, 2, 3, 4
FP64, FP32BULK = 0, 1
FORCE_NONE, FORCE_DNS, FORCE_GMRES, FORCE_KOLMO = 0, 1, 2, 3
_FLAV = {"ns2d": NS2D, "ns3d": NS3D, "euler": EULER3D, "euler3d": EULER3D, "bous": BOUS3D, "bous3d": BOUS3D}
_PREC = {"fp64": FP64, "fp32": FP32BULK, "fp32bulk": FP32BULK}
_FORCE = {"none": FORCE_NONE, "dns": FORCE_DNS, "gmres": FORCE_GMRES, "kolmogorov": FORCE_KOLMO,
          "kolmo": FORCE_KOLMO}


class PsgError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [("flavour", C.c_int), ("precision", C.c_int), ("batch", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("alpha", C.c_double),
                ("v2", C.c_double), ("dt", C.c_double), ("force", C.c_int), ("kf", C.c_double),
                ("ampfor", C.c_double), ("ri", C.c_double), ("sc", C.c_double), ("theta", C.c_double)]


class HGrid(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("alpha", C.c_double), ("ktx", C.c_int), ("kty", C.c_int), ("ktz", C.c_int),
                ("iktx", C.c_int), ("ikty", C.c_int), ("iktz", C.c_int), ("nx2", C.c_int),
                ("nkt", C.c_long), ("nr", C.c_long), ("n2", C.c_long)]


class NewtonOpts(C.Structure):
    _fields_ = [("max_newton", C.c_int), ("max_gmres", C.c_int), ("max_hook", C.c_int), ("netol", C.c_double),
                ("gmtol", C.c_double), ("eps", C.c_double), ("delta0", C.c_double), ("nshift", C.c_int),
                ("fixed_period", C.c_int), ("jv", C.c_int), ("time_policy", C.c_int), ("dt0", C.c_double),
                ("sy", C.c_double), ("stall_count", C.c_int), ("verbose", C.c_int), ("tl_record", C.c_int)]


class NewtonResult(C.Structure):
    _fields_ = [("T", C.c_double), ("sx", C.c_double), ("sz", C.c_double), ("ner", C.c_double),
                ("ner2", C.c_double), ("newton_iters", C.c_int), ("gmres_iters", C.c_int),
                ("hook_iters", C.c_int), ("status", C.c_int), ("seconds", C.c_double), ("phi_evals", C.c_long)]


class FloquetOpts(C.Structure):
    _fields_ = [("nev", C.c_int), ("ncv", C.c_int), ("max_restarts", C.c_int), ("tol", C.c_double),
                ("which", C.c_int), ("jv", C.c_int), ("eps", C.c_double), ("time_policy", C.c_int),
                ("dt0", C.c_double), ("sy", C.c_double), ("seed", C.c_ulonglong), ("verbose", C.c_int)]


class FloquetResult(C.Structure):
    _fields_ = [("nconv", C.c_int), ("matvecs", C.c_int), ("cycles", C.c_int), ("seconds", C.c_double)]


class RecOpts(C.Structure):
    """psg_rec_opts (include/psg_rec.h)."""
    _fields_ = [("rule", C.c_int), ("member", C.c_int), ("store_size", C.c_int), ("norm_width", C.c_int),
                ("nshift_x", C.c_int), ("nshift_y", C.c_int), ("nshift_z", C.c_int), ("nout", C.c_int),
                ("tstart", C.c_double), ("threshold", C.c_double), ("ri", C.c_double), ("keep_states", C.c_int)]


class RecGuess(C.Structure):
    _fields_ = [("index", C.c_int), ("store", C.c_int), ("time", C.c_double), ("v2", C.c_double),
                ("period", C.c_double), ("sx", C.c_double), ("sy", C.c_double), ("sz", C.c_double),
                ("residual", C.c_double), ("time_under", C.c_double)]


class RecCheck(C.Structure):
    _fields_ = [("checked", C.c_int), ("compared", C.c_int), ("store", C.c_int), ("count", C.c_int),
                ("residual", C.c_double), ("sx", C.c_double), ("sy", C.c_double), ("sz", C.c_double),
                ("time", C.c_double)]


NEWTON_STATUS = {0: "CONVERGED", 1: "GMRES FAILED", 2: "H-STEP FAILED", 3: "T-STEP FAILED", 4: "NEWTON STALLED",
                 5: "PERIOD TOO SMALL", 6: "NEWTON FAILED"}

_lib = None
_P, _D, _I, _L = C.c_void_p, C.c_double, C.c_int, C.c_long


class KV(C.Structure):
    """psgh_kv: labels and values of one parameter file (include/psg_host.h)."""
    _fields_ = [("n", C.c_int), ("key", (C.c_char * 40) * 48), ("kind", C.c_char * 48), ("ival", C.c_long * 48),
                ("fval", C.c_double * 48), ("sval", (C.c_char * 257) * 48)]


PARAMS_LAYOUTS = {"ns3d_dns": 0, "ns3d_gmres": 1, "euler": 2, "bous_dns": 3, "bous_gmres": 4, "ns2d_dns": 5,
                  "ns2d_gmres": 6, "extra_ns2d": 7, "extra_ns3d": 8, "extra_bous": 9}


def lib():
    """Load libpsg.so (raises PsgError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise PsgError("psgpu_amd: %s is not built; run `python -m psgpu_amd.build`" % LIBPATH)
    L = C.CDLL(LIBPATH)
    cfgp = C.POINTER(Config)
    L.psg_last_error.restype = C.c_char_p
    L.psg_sizes.argtypes = [cfgp] + [C.POINTER(C.c_long)] * 5
    L.psg_create.argtypes = [cfgp, C.POINTER(_P)]
    L.psg_destroy.argtypes = [_P]
    for f in ("psg_set_state", "psg_get_state", "psg_set_state_device", "psg_get_state_device"):
        getattr(L, f).argtypes = [_P, _I, _P]
    L.psg_state_ptr.argtypes = [_P, _I, C.POINTER(_P)]
    L.psg_run.argtypes = [_P, _I]
    L.psg_set_dt.argtypes = [_P, _D]
    L.psg_sync.argtypes = [_P]
    L.psg_diag.argtypes = [_P, _I, C.POINTER(_D), C.POINTER(_D), _P]
    L.psg_diag_noalloc.argtypes = [_P, _I, C.POINTER(_D), C.POINTER(_D), _P]
    L.psg_nonlin.argtypes = [_P, _I, _P, _P]
    L.psg_physical.argtypes = [_P, _I, _P]
    L.psg_timederiv.argtypes = [_P, _I, _P, _P]
    L.psg_set_tables.argtypes = [_P, _P, _P, _P]
    L.psg_stream.argtypes = [_P]
    L.psg_stream.restype = _P
    L.psg_profile_enable.argtypes = [_P, _I]
    L.psg_set_variant.argtypes = [_P, _I, _I]
    L.psg_set_graph.argtypes = [_P, _I]
    L.psg_get_variant.argtypes = [_P, _I]
    L.psg_autotune.argtypes = [_P, _I, C.c_char_p, _I]
    L.psg_profile_get.argtypes = [_P, _I, C.POINTER(C.c_char_p), C.POINTER(_D), C.POINTER(_L), C.POINTER(_I)]
    hg = C.POINTER(HGrid)
    L.psgh_grid3.argtypes = [hg, _I, _I, _I, _D]
    L.psgh_grid2.argtypes = [hg, _I, _I, _D]
    L.psgh_tables.argtypes = [hg, _P, _P, _P, _P]
    L.psgh_count_active.argtypes = [hg]
    L.psgh_count_active.restype = _L
    L.psgh_nutab.argtypes = [hg, _D, _D, _P, _P, _P]
    L.psgh_ikf_ikn.argtypes = [hg, _P, _P]
    L.psgh_twiddles.argtypes = [_I, _P]
    L.psgh_rng_seed.argtypes = [_I]
    L.psgh_ranno0.restype = _D
    L.psgh_ic_ns3d_ref.argtypes = [hg, _I, _D, _P, _P, _P]
    L.psgh_ic_ns2d_ref.argtypes = [hg, _I, _D, _P]
    L.psgh_ic_euler_ref.argtypes = [hg, _I, _D, _P, _P, _P]
    L.psgh_ic_bous_ref.argtypes = [hg, _I, _D, _P, _P, _P, _P]
    L.psgh_ic_solenoidal.argtypes = [hg, C.c_uint64, _D, _D, _P, _P, _P, _P]
    L.psgh_ic_ns2d_cfg.argtypes = [hg, C.c_uint64, _D, _P]
    L.psgh_params_read.argtypes = [C.c_char_p, _I, C.POINTER(KV), C.c_char_p, _I]
    L.psgh_zk_header.argtypes = [_P, _I, _I, _I, _I]
    L.psgh_zk_record.argtypes = [_P, _P, _I, _P, _L]
    L.psgh_rec_read.argtypes = [_P, _P, _L]
    L.psgh_rec_read.restype = _L
    L.psgh_raw_write.argtypes = [C.c_char_p, _P, _L]
    L.psgh_fortran_f.argtypes = [C.c_char_p, _I, _I]
    L.psgh_fortran_f.restype = _D
    L.psg_tangent_steps.argtypes = [_P, _P, _P, _I, _P, _P]
    L.psg_adjoint.argtypes = [_P, _P, _P, _I, _I, _P, _P]
    L.psg_residual_grad.argtypes = [_P, _P, _I, _I, _D, _D, _D, C.POINTER(_D), _P, _P]
    L.psg_svd.argtypes = [_I, _I, _P, _P, _P, _P]
    L.psg_hook_mu.argtypes = [_I, _P, _P, _D, _P]
    L.psg_hook_mu.restype = _D
    L.psg_hess_eig.argtypes = [_I, _P, _P, _P]
    L.psg_floquet.argtypes = [_P, _P, _D, _D, _D, _P, _P, _P, _P, _P, _P]
    L.psg_floquet_defaults.argtypes = [_P]
    L.psg_newton_defaults.argtypes = [_I, C.POINTER(NewtonOpts)]
    L.psg_newton_solve.argtypes = [_P, _P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_D), C.POINTER(NewtonOpts),
                                   C.POINTER(NewtonResult)]
    L.psg_flow_map.argtypes = [_P, _P, _D, _D, _I]
    L.psg_newton_batch.argtypes = [_P, _I, _P, _P, _P, _P, C.POINTER(NewtonOpts), _P]
    L.psg_stats_create.argtypes = [_P, _I, C.POINTER(_P)]
    L.psg_stats_update.argtypes = [_P, _L, _P]
    L.psg_stats_physical.argtypes = [_P, _I, _L, C.POINTER(_D), _P, _P, C.POINTER(_I)]
    L.psg_stats_spectra.argtypes = [_P, _P, _P, _I, C.POINTER(_I)]
    L.psg_stats_mean_fields.argtypes = [_P, _P]
    L.psg_stats_tavg.argtypes = [_P, _P]
    L.psg_stats_destroy.argtypes = [_P]
    L.psg_set_ampfor.argtypes = [_P, _D]
    L.psg_rec_defaults.argtypes = [_P, C.POINTER(RecOpts)]
    L.psg_rec_create.argtypes = [_P, C.POINTER(RecOpts), C.POINTER(_P)]
    L.psg_rec_update.argtypes = [_P, _L, _D, C.POINTER(RecGuess), _P, C.POINTER(RecCheck)]
    L.psg_rec_last.argtypes = [_P, _P, _P, C.POINTER(_I), _I]
    L.psg_rec_state.argtypes = [_P, _I, _P]
    L.psg_rec_destroy.argtypes = [_P]
    L.psg_tangent.argtypes = [_P, _P, _P, _D, _D, _I, _P, _P]
    L.psg_tangent_many.argtypes = [_P, _P, _I, _P, _D, _D, _I, _I, _P, _P]
    _lib = L
    return L


def last_error():
    return lib().psg_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise PsgError("%s failed (%d): %s" % (what, rc, last_error()))


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def device_count():
    return lib().psg_device_count()


# ------------------------------------------------------------------ host layer
def hgrid(flavour, nx, ny, nz=1, alpha=1.0):
    g = HGrid()
    if _FLAV.get(flavour, flavour) == NS2D:
        lib().psgh_grid2(C.byref(g), nx, ny, alpha)
    else:
        lib().psgh_grid3(C.byref(g), nx, ny, nz, alpha)
    return g


def tables(g):
    n = g.nkt
    kx, ky, kz = np.zeros(n), np.zeros(n), np.zeros(n)
    L = np.zeros(n, np.int32)
    lib().psgh_tables(C.byref(g), _ptr(kx), _ptr(ky), _ptr(kz), _ptr(L))
    return kx, ky, kz, L


def nutab(g, v2, dt):
    n = g.nkt
    nu, nuzn, nu1 = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().psgh_nutab(C.byref(g), v2, dt, _ptr(nu), _ptr(nuzn), _ptr(nu1))
    return nu, nuzn, nu1


def ic_ns3d_ref(g, iseed, ampv=1.0):
    s = np.zeros((3, g.nkt), np.complex128)
    lib().psgh_ic_ns3d_ref(C.byref(g), iseed, ampv, _ptr(s[0]), _ptr(s[1]), _ptr(s[2]))
    return s


def ic_euler_ref(g, iseed, ampv=1.0):
    """3D Euler IC (3D_Euler/src/init.F90:137-190): Taylor-Green pair plus a random kz=1 shell in zet."""
    s = np.zeros((3, g.nkt), np.complex128)
    lib().psgh_ic_euler_ref(C.byref(g), iseed, ampv, _ptr(s[0]), _ptr(s[1]), _ptr(s[2]))
    return s


def ic_bous_ref(g, iseed, ampv=1.0):
    """3D Boussinesq IC (3D_Boussinesq/src/init.F90:234-305): (xii, eta, zet, rho), rho = 0."""
    s = np.zeros((4, g.nkt), np.complex128)
    lib().psgh_ic_bous_ref(C.byref(g), iseed, ampv, _ptr(s[0]), _ptr(s[1]), _ptr(s[2]), _ptr(s[3]))
    return s


def ic_ns2d_ref(g, iseed, ampv=1.0):
    z = np.zeros((1, g.nkt), np.complex128)
    lib().psgh_ic_ns2d_ref(C.byref(g), iseed, ampv, _ptr(z[0]))
    return z


def ic_solenoidal(g, seed, energy=0.5, rho_scale=None):
    nf = 4 if rho_scale is not None else 3
    s = np.zeros((nf, g.nkt), np.complex128)
    rho = _ptr(s[3]) if rho_scale is not None else None
    lib().psgh_ic_solenoidal(C.byref(g), seed, energy, rho_scale or 0.0, _ptr(s[0]), _ptr(s[1]), _ptr(s[2]), rho)
    return s


def ic_ns2d_cfg(g, seed, ampv=1.0):
    z = np.zeros((1, g.nkt), np.complex128)
    lib().psgh_ic_ns2d_cfg(C.byref(g), seed, ampv, _ptr(z[0]))
    return z


def read_params(path, layout):
    """Read a reference parameter file with the reader of `layout` (PARAMS_LAYOUTS key);
    returns {label: value} in file order.  Raises ValueError on a label mismatch."""
    kv = KV()
    msg = C.create_string_buffer(256)
    rc = lib().psgh_params_read(str(path).encode(), PARAMS_LAYOUTS.get(layout, layout), C.byref(kv), msg, 256)
    if rc != 0:
        raise ValueError(msg.value.decode())
    out = {}
    for i in range(kv.n):
        k = kv.key[i].value.decode().rstrip(" =").rstrip()
        t = kv.kind[i:i + 1].decode()
        out[k] = {"I": lambda: kv.ival[i], "L": lambda: bool(kv.ival[i]), "F": lambda: kv.fval[i],
                  "A": lambda: kv.sval[i].value.decode()}[t]()
    return out


# ------------------------------------------------------------------ stepper
def make_config(flavour="ns3d", n=64, nx=None, ny=None, nz=None, alpha=1.0, precision="fp64", batch=1,
                v2=0.01, dt=0.002, force="kolmogorov", kf=4.0, ampfor=None, ri=0.0, sc=1.0, theta=np.pi / 2):
    fl = _FLAV.get(flavour, flavour)
    fo = _FORCE.get(force, force)
    ny = ny or n
    nz = nz or (1 if fl == NS2D else n)
    nx = nx or int(round(ny / alpha))
    if ampfor is None:
        ampfor = {FORCE_NONE: 0.0, FORCE_DNS: -0.25 * kf * kf, FORCE_GMRES: -0.25 * kf,
                  FORCE_KOLMO: -0.5 * kf}[fo]
    return Config(fl, _PREC.get(precision, precision), batch, nx, ny, nz, alpha, v2, dt, fo, kf, ampfor,
                  ri, sc, theta)


class Stepper:
    """Persistent GPU context for ``batch`` members of one flavour/grid."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else make_config(**kw)
        L = lib()
        sizes = [C.c_long() for _ in range(5)]
        _check(L.psg_sizes(C.byref(self.cfg), *[C.byref(s) for s in sizes]), "psg_sizes")
        self.iktx, self.ikty, self.iktz, self.nkt, self.nactive = (s.value for s in sizes)
        self.nf = {NS2D: 1, NS3D: 3, EULER3D: 3, BOUS3D: 4}[self.cfg.flavour]
        h = _P()
        _check(L.psg_create(C.byref(self.cfg), C.byref(h)), "psg_create")
        self.h = h

    @property
    def grid(self):
        """Host grid description (sizes and k tables) of this context."""
        c = self.cfg
        if c.flavour == NS2D:
            return hgrid(NS2D, c.nx, c.ny, 1, c.alpha)
        return hgrid(c.flavour, c.nx, c.ny, c.nz, c.alpha)

    @property
    def mask(self):
        """Active-mode mask L (float 0/1, length NKT)."""
        if not hasattr(self, "_mask"):
            self._mask = tables(self.grid)[3].astype(np.float64)
        return self._mask

    def close(self):
        if getattr(self, "h", None):
            lib().psg_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, state, member=0):
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt)
        _check(lib().psg_set_state(self.h, member, _ptr(s)), "psg_set_state")

    def get_state(self, member=0):
        s = np.zeros((self.nf, self.nkt), np.complex128)
        _check(lib().psg_get_state(self.h, member, _ptr(s)), "psg_get_state")
        return s

    def state_ptr(self, member=0):
        p = _P()
        _check(lib().psg_state_ptr(self.h, member, C.byref(p)), "psg_state_ptr")
        return p.value

    def run(self, nsteps):
        _check(lib().psg_run(self.h, int(nsteps)), "psg_run")

    def sync(self):
        _check(lib().psg_sync(self.h), "psg_sync")

    def set_dt(self, dt):
        _check(lib().psg_set_dt(self.h, dt), "psg_set_dt")
        self.cfg.dt = dt

    def diag(self, member=0):
        e, z = C.c_double(), C.c_double()
        kty = (self.ikty - 1) // 2
        spec = np.zeros(kty + 2)
        _check(lib().psg_diag(self.h, member, C.byref(e), C.byref(z), _ptr(spec)), "psg_diag")
        return e.value, z.value, spec

    def nonlin(self, state):
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt)
        out = np.zeros_like(s)
        _check(lib().psg_nonlin(self.h, 0, _ptr(s), _ptr(out)), "psg_nonlin")
        return out

    def timederiv(self, state):
        """dz/dt = -NU z + N(z) + F (masked) of one state (host arrays)."""
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt)
        out = np.zeros_like(s)
        _check(lib().psg_timederiv(self.h, 0, _ptr(s), _ptr(out)), "psg_timederiv")
        return out

    def physical(self, member=0):
        c = self.cfg
        nfo = 3 if c.flavour == NS2D else (7 if c.flavour == BOUS3D else 6)
        out = np.zeros((nfo, c.nz if c.flavour != NS2D else 1, c.ny, c.nx))
        _check(lib().psg_physical(self.h, member, _ptr(out)), "psg_physical")
        return out

    def flow_map(self, state, T, dt0, policy=0):
        """Phi_T(state) with the Newton time policy (0: dt=T/INT(T/dt0); 1: fixed dt0 + short last step)."""
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt).copy()
        _check(lib().psg_flow_map(self.h, _ptr(s), T, dt0, policy), "psg_flow_map")
        return s

    def tangent(self, state, dstate, T, dt0, policy=0):
        """(Phi_T(state), dPhi_T(state) dstate) from the tangent-linear stepper."""
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt)
        d = np.ascontiguousarray(dstate, dtype=np.complex128).reshape(self.nf, self.nkt)
        z1, dz1 = np.zeros_like(s), np.zeros_like(d)
        _check(lib().psg_tangent(self.h, _ptr(s), _ptr(d), T, dt0, policy, _ptr(z1), _ptr(dz1)), "psg_tangent")
        return z1, dz1

    def tangent_many(self, state, dstates, T, dt0, policy=0, fp32=False):
        """(Phi_T(state), [dPhi_T(state) d for d in dstates]) about one base (psg_tangent_many)."""
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt)
        d = np.ascontiguousarray(dstates, dtype=np.complex128).reshape(-1, self.nf, self.nkt)
        z1, dz1 = np.zeros_like(s), np.zeros_like(d)
        _check(lib().psg_tangent_many(self.h, _ptr(s), d.shape[0], _ptr(d), T, dt0, policy, int(bool(fp32)), _ptr(z1),
                                      _ptr(dz1)), "psg_tangent_many")
        return z1, dz1

    def tangent_steps(self, state, dstate, nsteps):
        """(Phi_n(state), dPhi_n(state) dstate) for n steps of dt (3D Navier-Stokes)."""
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt)
        d = np.ascontiguousarray(dstate, dtype=np.complex128).reshape(self.nf, self.nkt)
        z1, dz1 = np.zeros_like(s), np.zeros_like(d)
        _check(lib().psg_tangent_steps(self.h, _ptr(s), _ptr(d), nsteps, _ptr(z1), _ptr(dz1)), "psg_tangent_steps")
        return z1, dz1

    def adjoint(self, state, lam, nsteps, ckpt=20):
        """(Phi_n(state), dPhi_n(state)^T lam) with checkpoints every ckpt steps (3D Navier-Stokes)."""
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt)
        lt = np.ascontiguousarray(lam, dtype=np.complex128).reshape(self.nf, self.nkt)
        zT, l0 = np.zeros_like(s), np.zeros_like(lt)
        _check(lib().psg_adjoint(self.h, _ptr(s), _ptr(lt), nsteps, ckpt, _ptr(zT), _ptr(l0)), "psg_adjoint")
        return zT, l0

    def residual_grad(self, state, nsteps, ckpt=20, shift=(0.0, 0.0, 0.0)):
        """(J, dJ/dz0, Phi_n(z0)) for J = ||Phi_n(z0) - S z0||^2 (3D Navier-Stokes)."""
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(self.nf, self.nkt)
        g, zT, J = np.zeros_like(s), np.zeros_like(s), C.c_double()
        _check(lib().psg_residual_grad(self.h, _ptr(s), nsteps, ckpt, shift[0], shift[1], shift[2], C.byref(J),
                                       _ptr(g), _ptr(zT)), "psg_residual_grad")
        return J.value, g, zT

    def newton(self, guess, T, sx=0.0, sz=0.0, **opts):
        """Newton-GMRES-hookstep from (guess, T, sx[, sz]); returns (z, result dict)."""
        o = NewtonOpts()
        lib().psg_newton_defaults(self.cfg.flavour, C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        z = np.ascontiguousarray(guess, dtype=np.complex128).reshape(self.nf, self.nkt).copy()
        Tc, sxc, szc = C.c_double(T), C.c_double(sx), C.c_double(sz)
        r = NewtonResult()
        _check(lib().psg_newton_solve(self.h, _ptr(z), C.byref(Tc), C.byref(sxc), C.byref(szc), C.byref(o),
                                      C.byref(r)), "psg_newton_solve")
        res = {f: getattr(r, f) for f, _ in NewtonResult._fields_}
        res["status_name"] = NEWTON_STATUS.get(r.status, "?")
        return z, res

    def newton_batch(self, guesses, T, sx=None, **opts):
        """Newton-GMRES-hookstep on many guesses at once, the context's members serving as
        lanes (psg_newton_batch); returns (z[nguess, nf, nkt], T, sx, list of result dicts)."""
        o = NewtonOpts()
        lib().psg_newton_defaults(self.cfg.flavour, C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        z = np.ascontiguousarray(guesses, dtype=np.complex128).reshape(-1, self.nf, self.nkt).copy()
        ng = z.shape[0]
        Tv = np.ascontiguousarray(np.broadcast_to(np.asarray(T, float), (ng,))).copy()
        sv = np.zeros(ng) if sx is None else np.ascontiguousarray(np.broadcast_to(np.asarray(sx, float), (ng,))).copy()
        rs = (NewtonResult * ng)()
        _check(lib().psg_newton_batch(self.h, ng, _ptr(z), _ptr(Tv), _ptr(sv), None, C.byref(o),
                                      C.cast(rs, C.c_void_p)), "psg_newton_batch")
        out = []
        for r in rs:
            d = {f: getattr(r, f) for f, _ in NewtonResult._fields_}
            d["status_name"] = NEWTON_STATUS.get(r.status, "?")
            out.append(d)
        return z, Tv, sv, out

    def floquet(self, orbit, T, sx=0.0, sz=0.0, vectors=False, **opts):
        """Leading Floquet multipliers of S_s Phi_T at the orbit: (mu, err, vecs or None, result dict)."""
        o = FloquetOpts()
        lib().psg_floquet_defaults(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        z = np.ascontiguousarray(orbit, dtype=np.complex128).reshape(self.nf, self.nkt)
        mr, mi, er = np.zeros(o.nev), np.zeros(o.nev), np.zeros(o.nev)
        vecs = np.zeros((o.nev, self.nf, self.nkt), np.complex128) if vectors else None
        r = FloquetResult()
        _check(lib().psg_floquet(self.h, _ptr(z), T, sx, sz, C.byref(o), _ptr(mr), _ptr(mi), _ptr(er),
                                 _ptr(vecs) if vectors else None, C.byref(r)), "psg_floquet")
        return mr + 1j * mi, er, vecs, {f: getattr(r, f) for f, _ in FloquetResult._fields_}

    def set_graph(self, steps):
        """Steps per captured HIP graph in run() (0: direct launches only)."""
        _check(lib().psg_set_graph(self.h, steps), "psg_set_graph")

    def set_variant(self, variant, pass_=0):
        _check(lib().psg_set_variant(self.h, pass_, variant), "psg_set_variant")

    def get_variant(self, pass_=0):
        return lib().psg_get_variant(self.h, pass_)

    def autotune(self, steps=5):
        """Time every kernel-shape variant on this context's state, keep the fastest; returns the report."""
        buf = C.create_string_buffer(2048)
        _check(lib().psg_autotune(self.h, steps, buf, 2048), "psg_autotune")
        return buf.value.decode()

    def profile(self, on=True):
        _check(lib().psg_profile_enable(self.h, 1 if on else 0), "psg_profile_enable")

    def profile_report(self):
        rows = []
        i = 0
        while True:
            nm, ms, b, cnt = C.c_char_p(), C.c_double(), C.c_long(), C.c_int()
            if lib().psg_profile_get(self.h, i, C.byref(nm), C.byref(ms), C.byref(b), C.byref(cnt)) != 0:
                break
            rows.append({"kernel": nm.value.decode(), "avg_ms": ms.value, "bytes": b.value, "count": cnt.value})
            i += 1
        return rows


def _struct_dict(s):
    return {name: getattr(s, name) for name, _ in s._fields_}


class Recurrence:
    """Recurrence search over the DNS history of one context member
    (include/psg_rec.h; the reference's recurrence_check).  Call update(nt, dt)
    at every NOUT block of the time loop, after the steps up to it; it returns
    (guess dict or None, guess state or None, check dict)."""

    def __init__(self, stepper, **opts):
        self.st = stepper
        o = RecOpts()
        _check(lib().psg_rec_defaults(stepper.h, C.byref(o)), "psg_rec_defaults")
        for k, v in opts.items():
            if not hasattr(o, k):
                raise TypeError("unknown recurrence option %r" % k)
            setattr(o, k, v)
        self.opts = o
        self.ns = o.nshift_x * o.nshift_y * o.nshift_z
        h = _P()
        _check(lib().psg_rec_create(stepper.h, C.byref(o), C.byref(h)), "psg_rec_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().psg_rec_destroy(self.h)
            self.h = None

    __del__ = close

    def update(self, nt, dt, want_state=False):
        g, c = RecGuess(), RecCheck()
        gs = np.zeros((self.st.nf, self.st.nkt), np.complex128) if want_state else None
        r = lib().psg_rec_update(self.h, int(nt), float(dt), C.byref(g), _ptr(gs) if want_state else None,
                                 C.byref(c))
        _check(min(r, 0), "psg_rec_update")
        return (_struct_dict(g) if r == 1 else None), (gs if r == 1 else None), _struct_dict(c)

    def last(self):
        """Residual table of the last check, (compared stores x shifts), and the stores' slots."""
        cap = self.opts.store_size
        res = np.zeros(cap * self.ns)
        slots = np.zeros(cap, np.int32)
        nj = C.c_int()
        _check(lib().psg_rec_last(self.h, _ptr(res), _ptr(slots), C.byref(nj), cap), "psg_rec_last")
        return res[:nj.value * self.ns].reshape(nj.value, self.ns), slots[:nj.value]

    def state(self, store):
        s = np.zeros((self.st.nf, self.st.nkt), np.complex128)
        _check(lib().psg_rec_state(self.h, int(store), _ptr(s)), "psg_rec_state")
        return s


class Stats:
    """Run statistics of one context member on the GPU (include/psg_stats.h):
    computeStats sums with device time averages, and for the Boussinesq flavour
    the physical-space variances, max|u|, RiG plane sums, spectra and mean fields."""

    def __init__(self, stepper, member=0):
        self.st = stepper
        h = _P()
        _check(lib().psg_stats_create(stepper.h, member, C.byref(h)), "psg_stats_create")
        self.h = h
        self.nv = 12 if stepper.cfg.flavour == BOUS3D else 11

    def close(self):
        if getattr(self, "h", None):
            lib().psg_stats_destroy(self.h)
            self.h = None

    __del__ = close

    def update(self, count):
        v = np.zeros(self.nv)
        _check(lib().psg_stats_update(self.h, int(count), _ptr(v)), "psg_stats_update")
        return v

    def physical(self, what, count):
        c = self.st.cfg
        npl_max = max(c.ny, c.nz)
        na, sa = np.zeros(npl_max), np.zeros(npl_max)
        um, npl = C.c_double(0.0), C.c_int(0)
        _check(lib().psg_stats_physical(self.h, int(what), int(count), C.byref(um), _ptr(na), _ptr(sa),
                                        C.byref(npl)), "psg_stats_physical")
        return um.value, na[:npl.value], sa[:npl.value]

    def spectra(self):
        cap = self.st.cfg.nx + 1
        h, v = np.zeros(cap), np.zeros(cap)
        size = C.c_int(0)
        _check(lib().psg_stats_spectra(self.h, _ptr(h), _ptr(v), cap, C.byref(size)), "psg_stats_spectra")
        return h, v, size.value

    def mean_fields(self):
        """Boussinesq: rho, u, v, w, <v^2>, <w^2>, <u^2>-<u>^2; 3D Navier-Stokes: u, v, w, <v^2>, <u^2>-<u>^2."""
        c = self.st.cfg
        out = np.zeros((7 if c.flavour == BOUS3D else 5, c.nz, c.ny, c.nx))
        _check(lib().psg_stats_mean_fields(self.h, _ptr(out)), "psg_stats_mean_fields")
        return out

    def tavg(self):
        t = np.zeros((self.st.nf, self.st.nkt), np.complex128)
        _check(lib().psg_stats_tavg(self.h, _ptr(t)), "psg_stats_tavg")
        return t


STP_RMS, STP_UMAX, STP_RIG = 1, 2, 4
