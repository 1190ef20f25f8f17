"""The C Newton-GMRES driver (psgpu_amd/_lib/psg_gmres), a restatement of the
reference's src_gmres/main.F90 callers: parameter files, guess records in the
Zk.out layout, and the UPOinfo.dat / UPOs.out / Zk.out outputs
(2D_NavierStokes/src_RK4/src_gmres/main.F90:146-153, 224-225, 645-646,
1744-1835).  Argument and parameter errors are reported before any GPU work;
the GPU test converges the laminar Kolmogorov state from a perturbed guess."""
import os
import struct
import subprocess

import numpy as np
import pytest

import psgpu_amd as P

EXE = os.path.join(os.path.dirname(P.__file__), "_lib", "psg_gmres")

# order-fixed labels of 2D_NavierStokes/src_RK4/src_gmres/parameters.F90:146-203
MAIN_2D = ["ISEED", "IREST", "infile", "infile type", "Re", "TSTEP", "outfreq1", "outfreq2", "DELT", "AMPFOR",
           "ResidualThreshold", "STFLAG", "RCFLAG", "ICFLAG", "WNFLAG", "initial x shift", "initial y shift",
           "lamFLAG", "UPOflag"]
EXTRA_2D = ["IGUESSEND", "maxPERIOD", "minPERIOD", "gmres ResidualThreshold", "arclcFLAG", "1stOrder",
            "2ndOrder PredOnly", "Full 2ndOrder", "drN_reset", "Fixed Period", "Fixed Period Size", "delta_v2",
            "ArnoldiFLAG"]


def write_params(path, labels, values):
    with open(path, "w") as f:
        for lab in labels:
            f.write("%s = %s\n" % (lab, values[lab]))


def main_2d(**kw):
    v = {"ISEED": 1, "IREST": 1, "infile": "guess.out", "infile type": "D", "Re": "5.0", "TSTEP": "1.0",
         "outfreq1": "1.0", "outfreq2": "1.0", "DELT": "0.01", "AMPFOR": "-2.0", "ResidualThreshold": "1.0",
         "STFLAG": 0, "RCFLAG": 0, "ICFLAG": 0, "WNFLAG": 0, "initial x shift": "0.0", "initial y shift": "0.0",
         "lamFLAG": 0, "UPOflag": 1}
    v.update(kw)
    return v


def extra_2d(**kw):
    v = {"IGUESSEND": 1, "maxPERIOD": "100.0", "minPERIOD": "0.1", "gmres ResidualThreshold": "1.0",
         "arclcFLAG": "F", "1stOrder": "F", "2ndOrder PredOnly": "F", "Full 2ndOrder": "F", "drN_reset": "F",
         "Fixed Period": "T", "Fixed Period Size": "0.2", "delta_v2": "0.0", "ArnoldiFLAG": "F"}
    v.update(kw)
    return v


def rec(payload):
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def read_recs(path):
    out, b = [], open(path, "rb").read()
    i = 0
    while i < len(b):
        n = struct.unpack("<i", b[i:i + 4])[0]
        out.append(b[i + 4:i + 4 + n])
        i += n + 8
    return out


def setup_case(d, z, T=0.2, main=None, extra=None):
    g = P.hgrid("ns2d", 32, 32)
    write_params(os.path.join(d, "params_gmres.txt"), MAIN_2D, main or main_2d())
    write_params(os.path.join(d, "paramsEXTRA.txt"), EXTRA_2D, extra or extra_2d())
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<3i", 1, g.iktx, g.ikty)))
        f.write(rec(struct.pack("<6d", 0.0, 0.2, T, 0.0, 0.0, 0.3) + np.ascontiguousarray(z, np.complex128).tobytes()))
    return g


def run(args, d):
    return subprocess.run([EXE] + args + ["--dir", d], capture_output=True, text=True, timeout=300)


def test_driver_built_and_usage():
    assert os.path.exists(EXE)
    r = subprocess.run([EXE], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([EXE, "euler", "32"], capture_output=True, text=True)
    assert r.returncode == 1 and "flavour" in r.stderr


def test_driver_rejects_unsupported_inputs(tmp_path):
    d = str(tmp_path)
    g = P.hgrid("ns2d", 32, 32)
    r = run(["ns2d", "32", "--jv", "exact"], d)
    assert r.returncode == 1 and "--jv" in r.stderr
    setup_case(d, np.zeros(g.nkt, complex), main=main_2d(**{"infile type": "Z"}))
    r = run(["ns2d", "32"], d)
    assert r.returncode == 1 and "infile type unknown" in r.stderr
    setup_case(d, np.zeros(g.nkt, complex), extra=extra_2d(arclcFLAG="T", ArnoldiFLAG="T"))
    r = run(["ns2d", "32"], d)
    assert r.returncode == 1 and "exclusive" in r.stderr
    os.remove(os.path.join(d, "paramsEXTRA.txt"))
    r = run(["ns2d", "32"], d)
    assert r.returncode == 1


# order-fixed labels of 3D_Boussinesq/src/src_gmres/parameters.F90:140-160, 179-195
MAIN_BOUS_GMRES = ["ISEED", "IREST", "infile", "infile type", "KF", "Re", "Ri", "Theta", "Sc", "TSTEP",
                   "ResidualThreshold", "outfreq1", "outfreq2", "DELT", "AMPV", "OUTPLANE", "STATSFLAG", "RCFLAG",
                   "UPOflag", "ADAPTflag", "ISTOP"]
EXTRA_BOUS = ["IGUESSEND", "maxPERIOD", "gmres ResidualThreshold", "arclcFLAG", "1stOrder", "2ndOrder PredOnly",
              "Full 2ndOrder", "drN_reset", "Fixed Period", "Fixed Period Size", "delta_v2", "ArnoldiFLAG", "ParFLAG"]


def setup_bous(d, n, z, sc, **kw):
    """params.txt, paramsEXTRA.txt and a type-D guess record (header jw, IKTX, IKTY, IKTZ; time, v2, T,
    sx, sy, sz, NER, then xi, eta, zeta, rho; 3D_Boussinesq/src/src_gmres/init.F90:75-93)"""
    g = P.hgrid("bous", n, n, n)
    v = {"ISEED": 1, "IREST": 1, "infile": "guess.out", "infile type": "D", "KF": 4, "Re": "50.0", "Ri": "1.0",
         "Theta": "90.0", "Sc": "1.0", "TSTEP": "1.0", "ResidualThreshold": "1.0", "outfreq1": "1.0",
         "outfreq2": "1.0", "DELT": "0.01", "AMPV": "1.0", "OUTPLANE": 0, "STATSFLAG": 0, "RCFLAG": 0,
         "UPOflag": 1, "ADAPTflag": 0, "ISTOP": 0}
    e = {"IGUESSEND": 1, "maxPERIOD": "100.0", "gmres ResidualThreshold": "1.0", "arclcFLAG": "F", "1stOrder": "F",
         "2ndOrder PredOnly": "F", "Full 2ndOrder": "F", "drN_reset": "F", "Fixed Period": "F",
         "Fixed Period Size": "0.2", "delta_v2": "0.0", "ArnoldiFLAG": "F", "ParFLAG": 1}
    for k, x in kw.items():
        (e if k in e else v)[k] = x
    write_params(os.path.join(d, "params.txt"), MAIN_BOUS_GMRES, v)
    write_params(os.path.join(d, "paramsEXTRA.txt"), EXTRA_BOUS, e)
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<4i", 1, g.iktx, g.ikty, g.iktz)))
        f.write(rec(struct.pack("<7d", *sc) + np.ascontiguousarray(z, np.complex128).tobytes()))
    return g


def test_driver_bous_reads_type_d_guess_and_rejects_what_it_lacks(tmp_path):
    d = str(tmp_path)
    g = P.hgrid("bous", 16, 16, 16)
    # a guess period 1e-12 off a whole number of steps is snapped to it (main.F90:694-697)
    setup_bous(d, 16, np.zeros((4, g.nkt), complex), (1.5, 0.02, 0.3 + 1e-12, 0.25, 0.5, -0.125, 0.75))
    r = run(["bous", "16", "--list"], d)
    assert r.returncode == 0, r.stderr
    row = listed(r.stdout)[0]
    assert row == {"t": 1.5, "v2": 0.02, "T": 30 * 0.01, "sx": 0.25, "sy": 0.5, "sz": -0.125, "NER": 0.75}
    r = run(["bous", "16", "--jv", "tl"], d)
    assert r.returncode == 1 and "finite-difference" in r.stderr
    setup_bous(d, 16, np.zeros((4, g.nkt), complex), (0, 0.02, 0.3, 0, 0, 0, 1), arclcFLAG="T")
    r = run(["bous", "16"], d)
    assert r.returncode == 1 and "not built for the Boussinesq" in r.stderr

@pytest.mark.gpu
def test_driver_bous_arnoldi_flag_writes_floquet_rows(gpu, tmp_path):
    # ArnoldiFLAG on a Boussinesq guess: the 3D eigenvalues.dat layout (I3,1x,2F13.5 rows of
    # the unstable exponents, a blank row, then IREST, their sum and T; main.F90:968-983)
    import oracle_lib as O
    d, n = str(tmp_path), 16
    g = O.grid3(n, n, n)
    z0 = O.BOUS(n, v2=0.02, ri=1.0, theta=np.pi / 2, dt=0.01).run(O.ic_solenoidal(g, 75, 0.5, 0.2), 30)
    setup_bous(d, n, z0, (0.0, 0.02, 0.2, 0.0, 0.0, 0.0, 0.1), ArnoldiFLAG="T")
    r = run(["bous", str(n)], d)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = open(os.path.join(d, "eigenvalues.dat")).read().split("\n")
    blank = rows.index("     ")
    for row in rows[:blank]:
        i, lr, li = row.split()
        assert float(lr) > 0.0 and len(row) == 30
    last = rows[blank + 1].split()
    assert int(last[0]) == 1 and float(last[2]) == pytest.approx(0.2)
    assert float(last[1]) == pytest.approx(sum(float(x.split()[1]) for x in rows[:blank]), abs=1e-4)



def listed(out):
    rows = []
    for line in out.splitlines():
        if line.startswith("guess "):
            rows.append({k: float(v) for k, v in (f.split("=") for f in line.split(": ", 1)[1].split())})
    return rows


# (infile type, scalars ahead of the field) of 2D_NavierStokes/src_RK4/src_gmres/init.F90:66-106
# with t=1.5, v2=0.125, T=2.0, sx=0.25, sy=0.5, NER=0.75, NER2 or timeUNDER=0.875, RsymFLAG=0
REC2D = {"C": (1.5, 2.0, 0.25, 0.75, 0.875), "D": (1.5, 0.125, 2.0, 0.25, 0.75, 0.875),
         "E": (1.5, 9.0, 2.0, 0.25, 0.75, 0.875), "G": (1.5, 2.0, 0.25, 0.5, 0.75),
         "H": (1.5, 0.125, 2.0, 0.25, 0.5, 0.75), "I": (1.5, 9.0, 2.0, 0.25, 0.5, 0.75),
         "J": (1.5, 0.125, 2.0, 0.25, 0.5, 0.75, 0.875), "K": (1.5, 9.0, 2.0, 0.25, 0.5, 0.75, 0.875),
         "L": (1.5, 9.0, 2.0, 0.25, 0.5, 0.0, 0.75, 0.875)}


@pytest.mark.parametrize("itype", sorted(REC2D))
def test_driver_reads_every_2d_record_type(tmp_path, itype):
    # the scalars land where the reference's reader puts them: the y shift only from
    # the types that carry one (negated, main.F90:370), v2 only from D, H, J, and a
    # type-D record's NERin is not taken as a shift
    d = str(tmp_path)
    g = P.hgrid("ns2d", 32, 32)
    write_params(os.path.join(d, "params_gmres.txt"), MAIN_2D, main_2d(**{"infile type": itype, "Re": "5.0"}))
    write_params(os.path.join(d, "paramsEXTRA.txt"), EXTRA_2D, extra_2d(**{"Fixed Period": "F"}))
    sc = REC2D[itype]
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<3i", 1, g.iktx, g.ikty)))
        f.write(rec(struct.pack("<%dd" % len(sc), *sc) + np.zeros(g.nkt, np.complex128).tobytes()))
    r = run(["ns2d", "32", "--list"], d)
    assert r.returncode == 0, r.stderr
    (row,) = listed(r.stdout)
    has_y = itype in "GHIJKL"
    assert row["t"] == 1.5 and row["T"] == 2.0 and row["sx"] == 0.25
    assert row["sy"] == (-0.5 if has_y else 0.0)
    assert row["NER"] == 0.75
    assert row["v2"] == (0.125 if itype in "DHJ" else 0.2)


def test_driver_fixed_period_rescales_shift_and_snaps_dns_periods(tmp_path):
    d = str(tmp_path)
    g = P.hgrid("ns2d", 32, 32)
    write_params(os.path.join(d, "params_gmres.txt"), MAIN_2D, main_2d(**{"infile type": "H", "IREST": 1}))
    write_params(os.path.join(d, "paramsEXTRA.txt"), EXTRA_2D, extra_2d(**{"IGUESSEND": 2, "Fixed Period": "F"}))
    z = np.zeros(g.nkt, np.complex128).tobytes()
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<3i", 1, g.iktx, g.ikty)))
        f.write(rec(struct.pack("<6d", 0.0, 0.2, 1.23 + 3e-11, 0.1, 0.0, 0.5) + z))     # a DNS period, round-off
        f.write(rec(struct.pack("<6d", 0.0, 0.2, 1.2345, 0.1, 0.0, 0.5) + z))           # a Newton period
    rows = listed(run(["ns2d", "32", "--list"], d).stdout)
    assert rows[0]["T"] == round(1.23 / 0.01) * 0.01 and rows[1]["T"] == 1.2345
    write_params(os.path.join(d, "paramsEXTRA.txt"), EXTRA_2D, extra_2d(**{"Fixed Period": "T",
                                                                            "Fixed Period Size": "2.469"}))
    (row,) = listed(run(["ns2d", "32", "--list"], d).stdout)[:1]
    assert abs(row["sx"] - 2.469 * 0.1 / (1.23 + 3e-11)) < 1e-15 and row["T"] == 2.469


def test_driver_rejects_reflection_records_and_bad_headers(tmp_path):
    d = str(tmp_path)
    g = P.hgrid("ns2d", 32, 32)
    write_params(os.path.join(d, "params_gmres.txt"), MAIN_2D, main_2d(**{"infile type": "L"}))
    write_params(os.path.join(d, "paramsEXTRA.txt"), EXTRA_2D, extra_2d())
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<3i", 1, g.iktx, g.ikty)))
        f.write(rec(struct.pack("<8d", 0, 0, 1.0, 0, 0, 1.0, 0, 0) + np.zeros(g.nkt, np.complex128).tobytes()))
    r = run(["ns2d", "32", "--list"], d)
    assert r.returncode == 1 and "RsymFLAG" in r.stderr
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<3i", 1, g.iktx + 1, g.ikty)))
    r = run(["ns2d", "32", "--list"], d)
    assert r.returncode == 1 and "header" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("jv", ["fd", "tl32"])
def test_driver_converges_laminar_state_and_writes_outputs(gpu, tmp_path, jv):
    d = str(tmp_path)
    g = P.hgrid("ns2d", 32, 32)
    kx, ky, _, L = P.tables(g)
    lam = np.zeros(g.nkt, complex)
    lam[(kx == 0) & (ky == 4)] = -5.0 / 8          # laminar Kolmogorov vorticity at Re = 5
    z0 = lam + 1e-3 * P.ic_ns2d_cfg(g, 9, 1.0)[0]
    setup_case(d, z0)
    r = run(["ns2d", "32", "--noshift", "--maxnewton", "8", "--jv", jv], d)
    assert r.returncode == 0, r.stdout + r.stderr
    info = open(os.path.join(d, "UPOinfo.dat")).read().splitlines()
    assert len(info) == 2 and '"CONVERGED"' in info[1]
    upos = read_recs(os.path.join(d, "UPOs.out"))
    assert struct.unpack("<3i", upos[0]) == (1, g.iktx, g.ikty) and len(upos) == 2
    sc = struct.unpack("<6d", upos[1][:48])
    assert sc[2] == 0.2 and sc[5] <= 1e-12
    z = np.frombuffer(upos[1][48:], np.complex128)
    assert np.abs((z - lam)[L == 1]).max() < 1e-9
    zk = read_recs(os.path.join(d, "Zk.out"))
    assert len(zk) == 3 and np.array_equal(np.frombuffer(zk[1][48:], np.complex128), z0)


@pytest.mark.gpu
def test_driver_bous_two_shift_newton_outputs(gpu, tmp_path):
    # two Newton steps from a developed 16^3 state: the reference's UPOinfo.dat line
    # (format 1502), one Zk.out record of the last iterate (no header), and its
    # residual re-integrated on the oracle with both shifts applied on the host
    import oracle_lib as O
    from test_gpu_newton_bous import flow_host, ner_host
    d, n, dt0 = str(tmp_path), 16, 0.01
    g = O.grid3(n, n, n)
    z0 = O.BOUS(n, v2=0.02, ri=1.0, theta=np.pi / 2, dt=0.01).run(O.ic_solenoidal(g, 73, 0.5, 0.2), 30)
    setup_bous(d, n, z0, (0.0, 0.02, 0.3, 0.05, 0.0, -0.05, 0.3))
    r = run(["bous", str(n), "--maxnewton", "2"], d)
    assert r.returncode == 0, r.stdout + r.stderr
    info = open(os.path.join(d, "UPOinfo.dat")).read().splitlines()
    assert info[0].startswith("Guess No : Start time : guess Re :guess Period :guess Shiftx: guess Shifty: guess "
                              "Shiftz:")
    assert len(info) == 2
    line = info[1]
    assert line[:9] == "    1    " and line[9:25] == "    0.000000E+00" and line[25:41] == "    0.200000E-01"
    assert line[41:89] == "%12.5f%12.5f%12.5f%12.5f" % (0.3, 0.05, 0.0, -0.05)
    st = line[103:line.index(":", 106) + 2]
    assert st in (' :   "CONVERGED"   : ', ' : "GMRES FAILED"  : ', ' : "H-STEP FAILED" : ',
                  ' : "T-STEP FAILED" : ', ' : "NEWTON STALLED" : ', ' : "PERIOD TOO SMALL" : ',
                  ' : "NEWTON FAILED" : '), repr(st)
    zk = read_recs(os.path.join(d, "Zk.out"))
    assert len(zk) == 1 and len(zk[0]) == 9 * 8 + 4 * 16 * g.nkt
    sc = struct.unpack("<9d", zk[0][:72])
    assert sc[:4] == (0.0, 0.02, 1.0, 1.0) and sc[6] == 0.0
    z = np.frombuffer(zk[0][72:], np.complex128).reshape(4, g.nkt)
    T, sx, sz, ner = sc[4], sc[5], sc[7], sc[8]
    zT = flow_host(n, dict(v2=0.02, ri=1.0, theta=np.pi / 2), z, T, dt0)
    host = ner_host(g, z, zT, sx, sz)
    assert abs(host - ner) <= 1e-8 * ner, (host, ner)
    assert ner < 0.3
    upos = read_recs(os.path.join(d, "UPOs.out"))
    assert struct.unpack("<4i", upos[0]) == (1, g.iktx, g.ikty, g.iktz)


@pytest.mark.gpu
def test_driver_arnoldi_flag_writes_floquet_exponents(gpu, tmp_path):
    # ArnoldiFLAG: the guess (here the laminar Kolmogorov state at Re = 40, a fixed
    # point, unstable) is analysed instead of solved; eigenvalues.dat receives the
    # 2D layout of 2D_NavierStokes/src_RK4/src_gmres/main.F90:933-948
    d = str(tmp_path)
    g = P.hgrid("ns2d", 32, 32)
    kx, ky, _, L = P.tables(g)
    lam = np.zeros(g.nkt, complex)
    lam[(kx == 0) & (ky == 4)] = -2.0 / (16 / 40.0)
    write_params(os.path.join(d, "params_gmres.txt"), MAIN_2D, main_2d(Re="40.0"))
    write_params(os.path.join(d, "paramsEXTRA.txt"), EXTRA_2D, extra_2d(ArnoldiFLAG="T"))
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<3i", 1, g.iktx, g.ikty)))
        f.write(rec(struct.pack("<6d", 0.0, 1 / 40.0, 0.2, 0.0, 0.0, 0.0) + lam.tobytes()))
    r = run(["ns2d", "32", "--jv", "tl"], d)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = open(os.path.join(d, "eigenvalues.dat")).read().splitlines()
    assert lines[0].split() == ["UPO", "1"]
    rows = [tuple(float(x) for x in ln.split()) for ln in lines[1:]]
    assert rows[-1][0] == 40.0
    exps = [a for a, _ in rows[:-1]]
    assert exps and min(exps) > 0 and rows[-1][1] == max(exps)
    with P.Stepper(flavour="ns2d", n=32, v2=1 / 40.0, dt=0.01, force="kolmogorov", kf=4.0, ampfor=-2.0) as s:
        mu, err, _, res = s.floquet(lam[None], 0.2, nev=10, ncv=200, tol=1e-7, jv=1, dt0=0.01)
    lead = np.log(np.abs(mu[0])) / 0.2
    assert abs(rows[-1][1] - lead) <= 1e-4 * lead, (rows[-1], lead)


@pytest.mark.gpu
@pytest.mark.parametrize("jv", ["fd", "tl"])
def test_driver_arclength_continuation_traces_laminar_branch(gpu, tmp_path, jv):
    # arclcFLAG: from the laminar Kolmogorov state at Re = 5 (a fixed point, so the
    # period is fixed) the continuation steps Re by delta_v2, then follows the
    # branch z(v2) = -F/(v2 |k|^2) at (0, 4) by pseudo-arclength steps in (z, v2)
    # (2D_NavierStokes/src_RK4/src_gmres/main.F90:408-604, 1776-1803)
    d = str(tmp_path)
    g = P.hgrid("ns2d", 32, 32)
    kx, ky, _, L = P.tables(g)
    sel = (kx == 0) & (ky == 4)
    lam = np.zeros(g.nkt, complex)
    lam[sel] = -2.0 / (16 * 0.2)
    write_params(os.path.join(d, "params_gmres.txt"), MAIN_2D, main_2d())
    write_params(os.path.join(d, "paramsEXTRA.txt"), EXTRA_2D, extra_2d(arclcFLAG="T", IGUESSEND=4, delta_v2="0.02"))
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<3i", 1, g.iktx, g.ikty)))
        f.write(rec(struct.pack("<6d", 0.0, 0.2, 0.2, 0.0, 0.0, 0.0) + lam.tobytes()))
    r = run(["ns2d", "32", "--noshift", "--jv", jv], d)
    assert r.returncode == 0, r.stdout + r.stderr
    upos = read_recs(os.path.join(d, "UPOs.out"))[1:]
    assert len(upos) == 5, r.stdout
    pts = []
    for u in upos:
        sc = struct.unpack("<6d", u[:48])
        z = np.frombuffer(u[48:], np.complex128)
        v2 = sc[1]
        assert sc[5] <= 1e-12
        assert abs(z[sel][0] - (-2.0 / (16 * v2))) <= 1e-9 * abs(z[sel][0])     # on the laminar branch
        assert np.abs(z[~sel & (L == 1)]).max() <= 1e-9
        pts.append((v2, z[sel][0].real))
    v2s = [p[0] for p in pts]
    assert abs(v2s[0] - 0.18) < 1e-14 and all(b < a for a, b in zip(v2s, v2s[1:]))   # Re increases
    # arclength step: the second point lies twice the first secant's length along
    # that secant (main.F90:467-471 and the arclength row)
    t = np.array([pts[0][0] - 0.2, pts[0][1] - lam[sel][0].real])
    d0 = np.hypot(*t)
    proj = np.dot(np.array([pts[1][0] - pts[0][0], pts[1][1] - pts[0][1]]), t / d0)
    assert abs(proj - 2 * d0) <= 1e-6 * d0, (d0, proj)


# ---------------------------------------------------------------- DNS driver
DNS_EXE = os.path.join(os.path.dirname(P.__file__), "_lib", "psg_dns")
# order-fixed labels of 3D_NavierStokes/src/parameters.F90:134-149 and 3D_Euler/src/parameters.F90:127-136
MAIN_NS3D = ["ISEED", "IREST", "infile", "infile type", "KF", "Re", "TSTEP", "ResidualThreshold", "outfreq1",
             "outfreq2", "DELT", "AMPV", "STATSFLAG", "RCFLAG", "UPOflag", "ISTOP"]
MAIN_EULER = ["ISEED", "IREST", "infile", "infile type", "TSTEP", "outfreq1", "outfreq2", "DELT", "AMPV",
              "STATSFLAG"]


def ns3d_params(**kw):
    v = {"ISEED": 4242, "IREST": 0, "infile": "none", "infile type": "D", "KF": 4, "Re": "100.0", "TSTEP": "0.1",
         "ResidualThreshold": "1.0", "outfreq1": "0.025", "outfreq2": "0.05", "DELT": "0.005", "AMPV": "1.0",
         "STATSFLAG": 1, "RCFLAG": 0, "UPOflag": 0, "ISTOP": 0}
    v.update(kw)
    return v


def test_dns_driver_usage_and_parameter_errors(tmp_path):
    assert os.path.exists(DNS_EXE)
    r = subprocess.run([DNS_EXE], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([DNS_EXE, "ns1d", "32"], capture_output=True, text=True)
    assert r.returncode == 1 and "flavour" in r.stderr
    r = subprocess.run([DNS_EXE, "ns3d", "32", "--dir", str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 1   # no JOB0/params.txt
    os.makedirs(tmp_path / "JOB0")
    write_params(str(tmp_path / "JOB0" / "params.txt"), MAIN_NS3D[:5], ns3d_params())
    r = subprocess.run([DNS_EXE, "ns3d", "32", "--dir", str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 1 and "label" in r.stderr


@pytest.mark.gpu
def test_dns_driver_ns3d_reference_run_matches_oracle(gpu, tmp_path):
    import oracle_lib as O
    n = 32
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "JOB0"))
    write_params(os.path.join(d, "JOB0", "params.txt"), MAIN_NS3D, ns3d_params())
    r = subprocess.run([DNS_EXE, "ns3d", str(n), "--dir", d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    g = O.grid3(n, n, n)
    st = O.ic_ns3d_ref(g, 4242, 1.0)
    want = O.NS3D(n, v2=0.01, dt=0.005, force=1, kf=4.0).run(st, 20)   # NSTOP = INT(0.1/0.005)
    recs = read_recs(os.path.join(d, "JOB0", "final_000.out"))
    assert struct.unpack("<4i", recs[0]) == (1, g.iktx, g.ikty, g.iktz)
    got = np.frombuffer(recs[1][56:], np.complex128).reshape(3, g.nkt)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-11
    # every NOUT = 5 of 20 steps from NT = 5, and finalise.h's closing computeStats
    assert len(open(os.path.join(d, "JOB0", "stats.dat")).read().strip().splitlines()) == 4
    for f in ("avgs.dat", "means.dat", "pdfs.dat"):
        assert os.path.getsize(os.path.join(d, "JOB0", f)) > 0
    assert struct.unpack("<4i", read_recs(os.path.join(d, "JOB0", "Zk.out"))[0]) == (1, g.iktx, g.ikty, g.iktz)


@pytest.mark.gpu
@pytest.mark.parametrize("itype", ["A", "D"])
def test_dns_driver_ns3d_restart_uses_the_record_v2(gpu, tmp_path, itype):
    # 3D_NavierStokes/src/init.F90:79-125: types A (stream) and D (Zk.out layout) carry v2,
    # and the CN tables are built from it; params.txt says Re = 100, the record v2 = 1/40
    import oracle_lib as O
    n = 32
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "JOB0"))
    g = O.grid3(n, n, n)
    st = O.ic_solenoidal(g, 99, 0.5)
    body = struct.pack("<7d", 3.0, 1 / 40.0, 0.0, 0.0, 0.0, 0.0, 0.0) + st.astype(np.complex128).tobytes()
    with open(os.path.join(d, "JOB0", "guess.out"), "wb") as f:
        if itype == "A":
            f.write(body)
        else:
            f.write(rec(struct.pack("<4i", 1, g.iktx, g.ikty, g.iktz)) + rec(body))
    write_params(os.path.join(d, "JOB0", "params.txt"), MAIN_NS3D,
                 ns3d_params(IREST=1, ISTOP=1, infile="guess.out", **{"infile type": itype, "STATSFLAG": 0}))
    r = subprocess.run([DNS_EXE, "ns3d", str(n), "--dir", d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    want = O.NS3D(n, v2=1 / 40.0, dt=0.005, force=1, kf=4.0).run(st, 20)
    recs = read_recs(os.path.join(d, "JOB0", "final_001.out"))
    assert struct.unpack("<7d", recs[1][:56])[1] == 1 / 40.0
    got = np.frombuffer(recs[1][56:], np.complex128).reshape(3, g.nkt)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-11


@pytest.mark.gpu
def test_dns_driver_euler_reference_run_matches_oracle(gpu, tmp_path):
    import oracle_lib as O
    n = 32
    d = str(tmp_path)
    v = {"ISEED": 837920, "IREST": 0, "infile": "none", "infile type": "D", "TSTEP": "0.05", "outfreq1": "0.01",
         "outfreq2": "0.05", "DELT": "0.01", "AMPV": "0.001", "STATSFLAG": 0}
    write_params(os.path.join(d, "params.txt"), MAIN_EULER, v)
    r = subprocess.run([DNS_EXE, "euler", str(n), "--dir", d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    g = O.grid3(n, n, n)
    st = O.ic_euler_ref(g, 837920, 0.001)
    want = O.NS3D(n, dt=0.01, flavour=1, force=0).run(st, 5)   # NSTOP = INT(0.05/0.01)
    recs = read_recs(os.path.join(d, "final_000.out"))
    got = np.frombuffer(recs[1][56:], np.complex128).reshape(3, g.nkt)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-11


MAIN_BOUS = ["ISEED", "IREST", "infile", "infile type", "KF", "Re", "Ri", "Theta", "Sc", "TSTEP",
             "ResidualThreshold", "outfreq1", "outfreq2", "DELT", "AMPV", "STATSFLAG", "RCFLAG", "UPOflag",
             "ADAPTflag", "ISTOP", "Dtarget"]


@pytest.mark.gpu
def test_dns_driver_bous_reference_run_matches_oracle(gpu, tmp_path):
    import oracle_lib as O
    n = 32
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "JOB0"))
    v = {"ISEED": 5151, "IREST": 0, "infile": "none", "infile type": "D", "KF": 4, "Re": "100.0", "Ri": "1.0",
         "Theta": "90.0", "Sc": "1.0", "TSTEP": "0.06", "ResidualThreshold": "1.0", "outfreq1": "0.02",
         "outfreq2": "1.0", "DELT": "0.004", "AMPV": "1.0", "STATSFLAG": 0, "RCFLAG": 0, "UPOflag": 0,
         "ADAPTflag": 0, "ISTOP": 0, "Dtarget": "0.0"}
    write_params(os.path.join(d, "JOB0", "params.txt"), MAIN_BOUS, v)
    r = subprocess.run([DNS_EXE, "bous", str(n), "--dir", d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    g = O.grid3(n, n, n)
    st = O.ic_bous_ref(g, 5151, 1.0)
    # stepping while t <= TSTEP with dt = TSTEP / INT(TSTEP/DELT) (3D_Boussinesq/src/cuda.cu:68-99)
    dt, tstep = 0.004, 0.06
    nstop = int(tstep / dt)
    dd = tstep / nstop if nstop * dt != tstep else dt
    tme, nst = 0.0, 0
    while tme <= tstep:
        tme += dd
        nst += 1
    want = O.BOUS(n, v2=0.01, sc=1.0, ri=1.0, theta=np.pi / 2, dt=dd).run(st, nst)
    recs = read_recs(os.path.join(d, "JOB0", "final_000.out"))
    assert struct.unpack("<d", recs[1][:8])[0] == pytest.approx(tme, abs=1e-14)
    got = np.frombuffer(recs[1][56:], np.complex128).reshape(4, g.nkt)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-11


MAIN_NS2D_DNS = ["ISEED", "IREST", "infile", "infile type", "Re", "TSTEP", "outfreq1", "outfreq2", "DELT", "AMPFOR",
                 "SCALE", "ResidualThreshold", "STFLAG", "TDFCFLAG", "ICFLAG", "WNFLAG", "initial x shift",
                 "initial y shift", "lamFLAG", "UPOflag"]


def ns2d_dns_params(**kw):
    v = {"ISEED": 837920, "IREST": 0, "infile": "guess.out", "infile type": "D", "Re": "40.0", "TSTEP": "0.2",
         "outfreq1": "0.05", "outfreq2": "1.0", "DELT": "0.005", "AMPFOR": "-2.0", "SCALE": "0.0",
         "ResidualThreshold": "1.0", "STFLAG": 1, "TDFCFLAG": 0, "ICFLAG": 0, "WNFLAG": 0, "initial x shift": "0.0",
         "initial y shift": "0.0", "lamFLAG": 0, "UPOflag": 0}
    v.update(kw)
    return v


def test_dns_driver_ns2d_rejects_unknown_infile_type(tmp_path):
    d = str(tmp_path)
    write_params(os.path.join(d, "params.txt"), MAIN_NS2D_DNS, ns2d_dns_params(IREST=1, **{"infile type": "Q"}))
    r = subprocess.run([DNS_EXE, "ns2d", "32", "--dir", d], capture_output=True, text=True)
    assert r.returncode == 1 and "infile type" in r.stderr


@pytest.mark.gpu
def test_dns_driver_ns2d_reference_run_matches_oracle(gpu, tmp_path):
    import oracle_lib as O
    n = 64
    d = str(tmp_path)
    write_params(os.path.join(d, "params.txt"), MAIN_NS2D_DNS, ns2d_dns_params())
    r = subprocess.run([DNS_EXE, "ns2d", str(n), "--dir", d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    g = O.grid2(n, n)
    want = O.NS2D(n, v2=1 / 40.0, dt=0.005).run(O.ic_ns2d_ref(g, 837920, 1.0), 40)   # INT(0.2/0.005)
    recs = read_recs(os.path.join(d, "Zk.out"))
    assert struct.unpack("<3i", recs[0]) == (1, g.iktx, g.ikty) and len(recs) == 2
    sc = struct.unpack("<6d", recs[1][:48])
    assert sc[:3] == (0.2, 1 / 40.0, 0.2)          # TIME = TSTART = TSTEP, v2, TSTEP
    got = np.frombuffer(recs[1][48:], np.complex128)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-11
    assert len(open(os.path.join(d, "stats.dat")).read().strip().splitlines()) == 4   # every 10 of 40 steps


@pytest.mark.gpu
def test_dns_driver_ns2d_upo_restart_shift_and_residual(gpu, tmp_path):
    import oracle_lib as O
    n = 32
    d = str(tmp_path)
    g = O.grid2(n, n)
    kx, ky, _, L = O.tables(g)
    z0 = O.ic_ns2d_ref(g, 4321, 1.0)
    period, sx, v2 = 0.33, 0.7, 1 / 30.0
    with open(os.path.join(d, "guess.out"), "wb") as f:
        f.write(rec(struct.pack("<3i", 1, g.iktx, g.ikty)))
        f.write(rec(struct.pack("<6d", 5.0, v2, period, sx, 0.25, 0.0) + z0.astype(np.complex128).tobytes()))
    # Re in params.txt is overridden by the v2 of the type-D record (init.F90:82-84)
    write_params(os.path.join(d, "params.txt"), MAIN_NS2D_DNS,
                 ns2d_dns_params(IREST=1, UPOflag=1, STFLAG=0, Re="100.0", DELT="0.01"))
    r = subprocess.run([DNS_EXE, "ns2d", str(n), "--dir", d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    nit = int(period / 0.01)
    zT = O.NS2D(n, v2=v2, dt=period / nit).run(z0, nit)
    recs = read_recs(os.path.join(d, "Zk.out"))
    assert len(recs) == 3
    assert struct.unpack("<4d", recs[1][:32]) == (0.2, v2, period, sx)
    got = np.frombuffer(recs[1][48:], np.complex128)
    assert np.abs(got - zT).max() / np.abs(zT).max() < 1e-11
    shifted = np.where(L == 1, np.exp(-1j * kx * sx) * zT, 0)
    got_s = np.frombuffer(recs[2][48:], np.complex128)
    assert np.abs(got_s - shifted).max() / np.abs(shifted).max() < 1e-11
    ner = np.linalg.norm((shifted - z0)[L == 1]) / np.linalg.norm(zT[L == 1])
    line = [l for l in r.stdout.splitlines() if "||F|| / ||Z||" in l][0]
    assert float(line.split("=")[-1]) == pytest.approx(ner, rel=1e-9)
    assert "WARNING" in r.stdout   # NER differs from the record's 0.25
