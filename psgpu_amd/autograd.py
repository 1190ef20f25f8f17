"""torch.autograd bindings of the 3D Navier-Stokes flow map.

``FlowMap.apply(z, stepper, nsteps, ckpt)`` maps a real view of the spectral
vorticity, shape (3, NKT, 2) float64 (re, im of xii, eta, zet), to the real
view of Phi_n(z) computed by the HIP stepper; its backward is the HIP adjoint
(checkpoints every ``ckpt`` steps, recomputation per segment).  The real view
makes torch's gradient convention, dL/dre and dL/dim per stored mode, equal to
the adjoint in the inner product sum_L Re(conj(a) b); modes outside the active
mask L carry zero gradient (the map reads its input on L only).

PyTorch is only the autograd driver here: the forward and backward are the
C-ABI calls psg_run and psg_adjoint (include/psg_adjoint.h).
"""
import numpy as np
import torch

from . import Stepper


def _to_complex(t, nkt):
    a = t.detach().to("cpu", torch.float64).contiguous().numpy()
    return (a[..., 0] + 1j * a[..., 1]).reshape(3, nkt)


def _to_real(z, like):
    out = np.stack([z.real, z.imag], axis=-1)
    return torch.from_numpy(np.ascontiguousarray(out)).to(device=like.device, dtype=like.dtype)


class FlowMap(torch.autograd.Function):
    """z -> Phi_n(z) for a Navier-Stokes ``Stepper`` (flavour "ns3d")."""

    @staticmethod
    def forward(ctx, z, stepper: Stepper, nsteps: int, ckpt: int = 20):
        zc = _to_complex(z, stepper.nkt) * stepper.mask[None, :]
        stepper.set_state(zc)
        stepper.run(nsteps)
        zT = stepper.get_state()
        ctx.stepper, ctx.nsteps, ctx.ckpt = stepper, nsteps, ckpt
        ctx.save_for_backward(z)
        return _to_real(zT, z)

    @staticmethod
    def backward(ctx, gout):
        (z,) = ctx.saved_tensors
        s = ctx.stepper
        _, lam0 = s.adjoint(_to_complex(z, s.nkt), _to_complex(gout, s.nkt), ctx.nsteps, ctx.ckpt)
        return _to_real(lam0, z), None, None, None
