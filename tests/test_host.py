"""CPU tests of the product's C host layer and C-ABI library (no GPU calls):
the library loads and exports every declared symbol, and the host tables /
initial conditions agree bit-for-bit with the oracle's independent
restatement."""
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(psgh?_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


@pytest.mark.parametrize("header", ["psg.h", "psg_host.h", "psg_newton.h", "psg_adjoint.h"])
def test_library_exports_declared_symbols(header):
    L = P.lib()
    missing = [n for n in _declared(header) if not hasattr(L, n)]
    assert not missing, missing


def test_device_count_callable():
    assert P.device_count() >= 0


@pytest.mark.parametrize("dims", [(32, 32, 32, 1.0), (64, 32, 32, 0.5), (16, 16, 16, 1.0)])
def test_tables_match_oracle(dims):
    nx, ny, nz, a = dims
    g = P.hgrid("ns3d", nx, ny, nz, a)
    go = O.grid3(nx, ny, nz, a)
    assert (g.nkt, g.iktx, g.ikty, g.iktz) == (go.nkt, go.iktx, go.ikty, go.iktz)
    for x, y in zip(P.tables(g), O.tables(go)):
        assert np.array_equal(x, y)
    for x, y in zip(P.nutab(g, 0.01, 0.002), O.nutab(go, 0.01, 0.002)):
        assert np.array_equal(x, y)


def test_ics_match_oracle_bitwise():
    g, go = P.hgrid("ns3d", 32, 32, 32), O.grid3(32, 32, 32)
    assert np.array_equal(P.ic_ns3d_ref(g, 837920), O.ic_ns3d_ref(go, 837920))
    assert np.array_equal(P.ic_euler_ref(g, 837920, 0.3), O.ic_euler_ref(go, 837920, 0.3))
    assert np.array_equal(P.ic_bous_ref(g, 4242, 2.0), O.ic_bous_ref(go, 4242, 2.0))
    assert np.array_equal(P.ic_solenoidal(g, 99), O.ic_solenoidal(go, 99))
    assert np.array_equal(P.ic_solenoidal(g, 99, 0.5, 0.1), O.ic_solenoidal(go, 99, 0.5, 0.1))
    g2, go2 = P.hgrid("ns2d", 64, 64), O.grid2(64, 64)
    assert np.array_equal(P.ic_ns2d_ref(g2, 1234)[0], O.ic_ns2d_ref(go2, 1234))
    assert np.array_equal(P.ic_ns2d_cfg(g2, 77)[0], O.ic_ns2d_cfg(go2, 77))


def test_active_counts():
    assert P.lib().psgh_count_active(P.hgrid("ns2d", 256, 256)) == 11214
    assert P.lib().psgh_count_active(P.hgrid("ns3d", 128, 64, 64, 0.5)) == 36109


def test_fortran_f_descriptor():
    f = lambda s: P.lib().psgh_fortran_f(s.encode(), 12, 4)
    assert f("20.0") == 20.0
    assert f("  0.005") == 0.005
    assert f("20") == 0.002           # F12.4 without a decimal point: last 4 digits are the fraction
    assert f("1.5e2") == 150.0
    assert f("") == 0.0


def test_create_fails_loudly_without_device():
    if P.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(P.PsgError):
        P.Stepper(flavour="ns3d", n=32)


@pytest.mark.parametrize("flavour", ["ns3d_dns", "ns3d_gmres", "euler", "ns2d_gmres", "ns2d_dns", "bous_dns",
                                     "bous_gmres"])
def test_legacy_libraries_export_reference_symbols(flavour):
    import ctypes
    path = os.path.join(ROOT, "psgpu_amd", "_lib", "libpsg_legacy_%s.so" % flavour)
    lib = ctypes.CDLL(path)
    declared = _declared("psg_legacy.h")
    assert "timestep_cuda_%s" % flavour in open(os.path.join(ROOT, "include", "psg_legacy.h")).read()
    for sym in ("timestep_cuda_", "set_gpu_", "timestep_cuda_%s" % flavour):
        assert hasattr(lib, sym), sym
    assert declared == []   # the legacy header declares no psg_* symbols
    if flavour == "bous_gmres":
        assert hasattr(lib, "timederiv_cuda_")
