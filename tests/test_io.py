"""CPU tests of the drop-in file formats: the parameter readers on the
reference's own shipped parameter files (tests/golden/reference_files, copied
by tests/golden/copy_reference_files.sh), and the Fortran unformatted records
of Zk.out / UPOs.out against the reference's shipped Zk.out header."""
import ctypes as C
import os

import numpy as np
import pytest

import psgpu_amd as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_files")
libc = C.CDLL(None)
libc.fopen.restype = C.c_void_p
libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
libc.fclose.argtypes = [C.c_void_p]


def test_params_ns3d_dns():
    p = P.read_params(os.path.join(GOLD, "ns3d_params.txt"), "ns3d_dns")
    assert p == {"ISEED": 837920, "IREST": 0, "infile": "none", "infile type": "D", "KF": 4, "Re": 20.0,
                 "TSTEP": 50.0, "ResidualThreshold": 1.0, "outfreq1": 0.1, "outfreq2": 10.0, "DELT": 0.01,
                 "AMPV": 1.0, "STATSFLAG": 1, "RCFLAG": 0, "UPOflag": 0, "ISTOP": 1}


def test_params_bous_dns_and_extra():
    p = P.read_params(os.path.join(GOLD, "bous_params.txt"), "bous_dns")
    assert (p["Re"], p["Ri"], p["Theta"], p["Sc"], p["ADAPTflag"], p["Dtarget"]) == (50.0, 1.0, 90.0, 1.0, 1, 0.0)
    assert list(p)[:5] == ["ISEED", "IREST", "infile", "infile type", "KF"]
    for name, lay in (("bous_paramsEXTRA.txt", "extra_bous"), ("ns3d_paramsEXTRA.txt", "extra_ns3d")):
        e = P.read_params(os.path.join(GOLD, name), lay)
        assert e["IGUESSEND"] == 53 and e["maxPERIOD"] == 25.0 and e["gmres ResidualThreshold"] == 0.6
        assert e["arclcFLAG"] is False and e["Full 2ndOrder"] is True and e["Fixed Period Size"] == 0.5


def test_params_euler_exponent_field():
    p = P.read_params(os.path.join(GOLD, "euler_params.txt"), "euler")
    assert p["AMPV"] == 1e-3 and p["outfreq1"] == 0.01 and p["STATSFLAG"] == 1


def test_params_label_mismatch_is_an_error():
    # the Euler file read with the Navier-Stokes table stops at the first differing label
    with pytest.raises(ValueError, match="bad label"):
        P.read_params(os.path.join(GOLD, "euler_params.txt"), "ns3d_dns")


def test_zk_header_matches_reference_file(tmp_path):
    path = str(tmp_path / "Zk.out").encode()
    f = libc.fopen(path, b"wb")
    assert P.lib().psgh_zk_header(f, 1, 86, 171, 171) == 0
    libc.fclose(f)
    assert open(path, "rb").read() == open(os.path.join(GOLD, "euler_Zk.out"), "rb").read()


def test_zk_record_round_trip(tmp_path):
    path = str(tmp_path / "UPOs.out").encode()
    rng = np.random.default_rng(3)
    fields = rng.standard_normal(2 * 3 * 500)
    scal = np.array([1.5, 0.01, 7.25, 0.3, 0.0, 0.1, 1e-11])     # t, v2, T, sx, sy, sz, NER
    f = libc.fopen(path, b"wb")
    assert P.lib().psgh_zk_header(f, 1, 10, 25, 2) == 0
    assert P.lib().psgh_zk_record(f, scal.ctypes.data, 7, fields.ctypes.data, 3 * 500) == 0
    libc.fclose(f)
    raw = open(path, "rb").read()
    n = 7 * 8 + fields.nbytes
    # gfortran sequential record: 4-byte length, payload, 4-byte length
    off = 24
    assert np.frombuffer(raw[off:off + 4], np.int32)[0] == n
    pay = np.frombuffer(raw[off + 4:off + 4 + n], np.float64)
    assert np.array_equal(pay[:7], scal) and np.array_equal(pay[7:], fields)
    assert np.frombuffer(raw[off + 4 + n:off + 8 + n], np.int32)[0] == n
    f = libc.fopen(path, b"rb")
    buf = np.zeros(n // 8 + 4)
    assert P.lib().psgh_rec_read(f, buf.ctypes.data, buf.nbytes) == 16
    assert P.lib().psgh_rec_read(f, buf.ctypes.data, buf.nbytes) == n
    libc.fclose(f)
    assert np.array_equal(buf[:7], scal)
