"""GPU tests of the 3D Navier-Stokes tangent linear and adjoint (checkpointed
backward sweep) and of the torch.autograd harness around them.

Pins: the tangent linear against central differences of the FP64 oracle's
flow map; the adjoint against the tangent linear by the dot-product identity
<J v, w> = <v, J^T w>; the gradient of J = ||Phi_n(z) - S z||^2 against
central differences (epsilon 1e-6) of the oracle's J along seeded solenoidal
directions.  Inner product: sum over active modes of Re(conj(a) b)."""
import numpy as np
import pytest

import oracle_lib as O
import psgpu_amd as P

pytestmark = pytest.mark.gpu

N, V2, DT = 32, 0.02, 0.004


def ip(a, b, L):
    return float(np.sum((np.conj(a) * b).real * L[None, :]))


def direction(g, seed):
    d = O.ic_solenoidal(g, seed, 0.5)          # solenoidal vorticity on the active modes
    return d / np.sqrt(np.sum(np.abs(d) ** 2))


def base_state(g):
    orc = O.NS3D(N, v2=V2, dt=DT, force=3)
    return orc.run(O.ic_solenoidal(g, 5, 0.5), 25)   # a state with a developed cascade


def stepper(prec="fp64"):
    return P.Stepper(flavour="ns3d", n=N, v2=V2, dt=DT, force="kolmogorov", precision=prec)


def test_tl_3d_matches_oracle_central_fd(gpu):
    g = O.grid3(N, N, N)
    z, d = base_state(g), direction(g, 11)
    nst, eps = 20, 1e-6
    orc = O.NS3D(N, v2=V2, dt=DT, force=3)
    fd = (orc.run(z + eps * d, nst) - orc.run(z - eps * d, nst)) / (2 * eps)
    with stepper() as s:
        zT, dzT = s.tangent_steps(z, d, nst)
    assert np.abs(zT - orc.run(z, nst)).max() < 1e-11 * np.abs(zT).max()
    err = np.abs(dzT - fd).max() / np.abs(fd).max()
    assert err < 1e-7, err


def test_tl_3d_through_newton_tangent(gpu):
    # psg_tangent (Newton's Jv) on the 3D flavour agrees with psg_tangent_steps
    g = O.grid3(N, N, N)
    z, d = base_state(g), direction(g, 12)
    with stepper() as s:
        _, a = s.tangent_steps(z, d, 10)
        _, b = s.tangent(z, d, 10 * DT, DT, 1)
    assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("ckpt", [1, 7, 30])
def test_adjoint_dot_product(gpu, ckpt):
    g = O.grid3(N, N, N)
    L = O.tables(g)[3].astype(float)
    z, v, w = base_state(g), direction(g, 21), direction(g, 22)
    nst = 30
    with stepper() as s:
        _, jv = s.tangent_steps(z, v, nst)
        zT, jtw = s.adjoint(z, w, nst, ckpt)
        after = s.get_state()
    lhs, rhs = ip(jv, w, L), ip(v, jtw, L)
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)
    assert np.abs(after - zT).max() == 0.0         # the context holds Phi_n(z) afterwards


def test_residual_gradient_vs_oracle_fd(gpu):
    g = O.grid3(N, N, N)
    L = O.tables(g)[3].astype(float)
    z = base_state(g)
    nst, eps = 40, 1e-6
    orc = O.NS3D(N, v2=V2, dt=DT, force=3)
    kx, ky, kz, _ = O.tables(g)
    for shift in [(0.0, 0.0, 0.0), (0.3, 0.0, -0.2)]:
        S = np.exp(-1j * (kx * shift[0] + ky * shift[1] + kz * shift[2]))[None, :]

        def J(x):
            r = (orc.run(x, nst) - S * x) * L[None, :]
            return float(np.sum(np.abs(r) ** 2))

        with stepper() as s:
            Jg, grad, _ = s.residual_grad(z, nst, 9, shift)
        assert abs(Jg - J(z)) <= 1e-11 * J(z)
        for k in range(4):
            v = direction(g, 100 + k)
            fd = (J(z + eps * v) - J(z - eps * v)) / (2 * eps)
            ad = ip(grad, v, L)
            assert abs(ad - fd) <= 1e-6 * abs(fd), (shift, k, ad, fd)


def test_adjoint_fp32_bulk_error_stated(gpu):
    # FP32-bulk transforms: the gradient is compared with the FP64 adjoint;
    # the bound is the stated FP32-bulk figure for this case
    g = O.grid3(N, N, N)
    z = base_state(g)
    with stepper() as s:
        _, g64, _ = s.residual_grad(z, 20, 10)
    with stepper("fp32") as s:
        _, g32, _ = s.residual_grad(z, 20, 10)
    err = np.abs(g32 - g64).max() / np.abs(g64).max()
    assert err < 1e-4, err


def test_autograd_flow_map(gpu):
    torch = pytest.importorskip("torch")
    from psgpu_amd.autograd import FlowMap
    g = O.grid3(N, N, N)
    L = O.tables(g)[3].astype(float)
    z = base_state(g) * L[None, :]
    nst = 15
    with stepper() as s:
        zr = torch.tensor(np.stack([z.real, z.imag], -1), dtype=torch.float64, requires_grad=True)
        out = FlowMap.apply(zr, s, nst, 4)
        loss = ((out - zr) ** 2).sum()
        loss.backward()
        gt = zr.grad.numpy()
        gt = gt[..., 0] + 1j * gt[..., 1]
        Jg, grad, _ = s.residual_grad(z, nst, 4)
    assert abs(loss.item() - Jg) <= 1e-12 * Jg
    assert np.abs(gt - grad).max() <= 1e-12 * np.abs(grad).max()
    orc = O.NS3D(N, v2=V2, dt=DT, force=3)
    v, eps = direction(g, 7), 1e-6
    Jf = lambda x: float(np.sum(np.abs((orc.run(x, nst) - x) * L[None, :]) ** 2))
    fd = (Jf(z + eps * v) - Jf(z - eps * v)) / (2 * eps)
    assert abs(ip(gt, v, L) - fd) <= 1e-6 * abs(fd)
