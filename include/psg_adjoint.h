/*
 * Tangent linear and adjoint of the 3D Navier-Stokes flow map (Heun + CN,
 * 3D_NavierStokes/src/kernels.h:309-464) on the GPU.  No counterpart in the
 * reference: its Newton solver differentiates the flow map by finite
 * differences (3D_NavierStokes/src_gmres/main.F90:1096-1107, epsilon 1e-7).
 *
 * All arrays are host arrays of 3*NKT complex doubles (xii, eta, zet) in the
 * psg.h storage order; member 0 of the context is used, with the context's
 * dt, forcing and CN tables.  Inner product: <a,b> = sum over the active
 * modes L of Re(conj(a) b), summed over the three fields; initial states are
 * taken on the active modes (masked on entry).
 */
#ifndef PSG_ADJOINT_H
#define PSG_ADJOINT_H
#include "psg.h"
#ifdef __cplusplus
extern "C" {
#endif

/* (zT, dzT) = (Phi_n(z0), dPhi_n(z0) dz0) for n steps: base and perturbation
 * stepped together.  zT may be NULL. */
int psg_tangent_steps(psg_ctx *ctx, const double *z0, const double *dz0, int nsteps, double *zT, double *dzT);

/* lam0 = dPhi_n(z0)^T lamT, with a checkpoint of the forward trajectory every
 * ckpt steps and recomputation of each segment during the backward sweep
 * (device memory: (ceil(n/ckpt) + 2 ckpt + 5) states).  zT (may be NULL)
 * receives Phi_n(z0); afterwards the context holds Phi_n(z0). */
int psg_adjoint(psg_ctx *ctx, const double *z0, const double *lamT, int nsteps, int ckpt, double *zT, double *lam0);

/* J = ||Phi_n(z0) - S z0||^2 and its gradient with respect to z0 on the
 * active modes, S z = exp(-i (kx sx + ky sy + kz sz)) z the shift of the
 * Newton solver: grad = dPhi^T (2r) - S^{-1} (2r), r = Phi_n(z0) - S z0.
 * zT (may be NULL) receives Phi_n(z0). */
int psg_residual_grad(psg_ctx *ctx, const double *z0, int nsteps, int ckpt, double sx, double sy, double sz,
                      double *J, double *grad, double *zT);

#ifdef __cplusplus
}
#endif
#endif
