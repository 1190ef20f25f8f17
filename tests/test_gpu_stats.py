"""GPU run statistics (include/psg_stats.h) against the oracle (oracle/or_stats.c)
on seeded states: computeStats sums and the device time averages (3D
Navier-Stokes and Boussinesq, both gravity orientations), the Boussinesq
physical-space pass (velocity variances, max|u|, RiG plane sums) against the
oracle's physical fields, the spectra and the mean fields."""
import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P

pytestmark = pytest.mark.gpu


def close(a, b, rtol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1e-300, float(np.max(np.abs(b))))
    return float(np.max(np.abs(a - b))) <= rtol * scale


def states(g, nf, n, seed):
    rng = np.random.default_rng(seed)
    kx, ky, kz, L = O.tables(g)
    return [(rng.standard_normal((nf, g.nkt)) + 1j * rng.standard_normal((nf, g.nkt))) * L * 0.1
            for _ in range(n)]


@pytest.mark.parametrize("n", [32, 128])
def test_ns3d_sums_and_averages(gpu, n):
    g = O.grid3(n, n, n)
    st = P.Stepper(flavour="ns3d", n=n, force="dns", kf=4.0)
    s = P.Stats(st)
    t = np.zeros((3, g.nkt), np.complex128)
    for c, z in enumerate(states(g, 3, 3, 5)):
        # the random states hold energy at the forcing modes too: the input sum is non-zero
        st.set_state(z)
        v = s.update(c)
        w = O.stats_ns3d(g, z, t, c, st.cfg.ampfor, 4.0)
        for k in range(11):
            assert close(v[k], w[k]), (c, k, v[k], w[k])
    assert close(s.tavg(), t, 1e-14)
    s.close()
    st.close()


@pytest.mark.parametrize("theta", [np.pi / 2, np.pi / 8])
def test_bous_sums_physical_spectra_means(gpu, theta):
    n, ri = 32, 0.7
    g = O.grid3(n, n, n)
    st = P.Stepper(flavour="bous", n=n, ri=ri, theta=theta, v2=0.01, dt=0.002)
    s = P.Stats(st)
    t = np.zeros((4, g.nkt), np.complex128)
    ns = O.NS3D(n)
    zero = np.zeros(g.nkt, np.complex128)
    rms = np.zeros((3, n, n, n))
    for c, z in enumerate(states(g, 4, 3, 7)):
        st.set_state(z)
        v = s.update(c)
        w = O.stats_bous(g, z, t, c, st.cfg.ampfor, st.cfg.kf, ri, theta)
        for k in range(12):
            assert close(v[k], w[k]), (c, k, v[k], w[k])
        um, na, sa = s.physical(P.STP_RMS | P.STP_UMAX | P.STP_RIG, c)
        ph = ns.physical(z[:3])
        ro = ns.physical(np.stack([z[3], zero, zero]))[0]
        U, V, W = ph[3], ph[4], ph[5]
        assert close(um, np.abs(U).max())
        ona, osa = O.rig(n, n, n, theta, U, V, W, ro)
        assert len(na) == len(ona)
        assert close(na, ona, 1e-11) and close(sa, osa, 1e-11)
        rms = (rms * c + np.stack([U, V, W]) ** 2) / (c + 1.0)
    assert close(s.tavg(), t, 1e-14)
    h, v, size = s.spectra()
    oh, ov, osize = O.spectra(g, t, theta)
    assert size == osize and close(h, oh) and close(v, ov)
    m = s.mean_fields()
    pt = ns.physical(t[:3])
    rt = ns.physical(np.stack([t[3], zero, zero]))[0]
    assert close(m[0], rt) and close(m[1], pt[3]) and close(m[2], pt[4]) and close(m[3], pt[5])
    assert close(m[4], rms[1]) and close(m[5], rms[2])
    assert close(m[6], rms[0] - pt[3] ** 2, 1e-11)
    s.close()
    st.close()


def test_set_ampfor_changes_the_forced_step(gpu):
    # the Dtarget throttle's entry point: a new amplitude reaches the captured steps
    n = 32
    g = O.grid3(n, n, n)
    z = states(g, 4, 1, 9)[0]
    st = P.Stepper(flavour="bous", n=n, ri=1.0, theta=np.pi / 2, v2=0.01, dt=0.002)
    st.set_state(z)
    st.run(20)
    P.lib().psg_set_ampfor(st.h, -3.0)
    st.run(20)
    got = st.get_state()
    o1 = O.BOUS(n, v2=0.01, ri=1.0, theta=np.pi / 2, dt=0.002, ampfor=-2.0)
    o2 = O.BOUS(n, v2=0.01, ri=1.0, theta=np.pi / 2, dt=0.002, ampfor=-3.0)
    want = o2.run(o1.run(z, 20), 20)
    assert close(got, want, 1e-11)
    st.close()
