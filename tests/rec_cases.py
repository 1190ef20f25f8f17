"""Synthetic DNS histories with a planted (relative) periodic episode, for the
recurrence-search tests.

A history entry n (time t_n = n h) is
    z(t) = exp(i kx a t) u(t),
    u(t) = A + B cos(2 pi t / P) + C sin(2 pi t / P) + eps t N          (t <= t_break)
           ... + g (t - t_break) G                                      (t >  t_break)
with A, B, C, N, G random masked states.  Shifting z(t) by s = a P in x,
exp(-i kx s) z(t) differs from z(t - P) by exactly eps P N (times a phase), so
the search's residual at lag P is eps P |N| / |z(t)| on the norm subset, with
the shift index `sx_index` of the reference's grid 2 pi i / (NSHIFTX alpha);
after t_break the state runs away and every residual exceeds the threshold,
which ends the output sequence.
"""
import numpy as np


def masked_random(rng, nf, L, scale=1.0):
    z = rng.standard_normal((nf, L.size)) + 1j * rng.standard_normal((nf, L.size))
    return scale * z * L[None, :]


def wave_history(kx, L, nf, n_entries, h, P, sx_index, nshift_x, alpha=1.0, eps=1e-3, t_break=None, g=10.0,
                 seed=0):
    rng = np.random.default_rng(seed)
    A = masked_random(rng, nf, L)
    B = masked_random(rng, nf, L, 0.5)
    Cc = masked_random(rng, nf, L, 0.5)
    N = masked_random(rng, nf, L)
    G = masked_random(rng, nf, L)
    a = (2 * np.pi * sx_index / (nshift_x * alpha)) / P
    out = np.empty((n_entries, nf, L.size), np.complex128)
    for n in range(n_entries):
        t = n * h
        u = A + B * np.cos(2 * np.pi * t / P) + Cc * np.sin(2 * np.pi * t / P) + (eps / P) * t * N
        if t_break is not None and t > t_break:
            u = u + g * (t - t_break) * G
        out[n] = np.exp(1j * kx * a * t)[None, :] * u
    return out, dict(a=a, N=N, eps=eps / P)


def subset_index(g, W):
    """The norm subset in the reference's order (0 <= ikx < W; 3D: ikz, iky from
    KTZ-W to KTZ+W-1; 2D: iky from KTY-W to KTY+W-1)."""
    if g.dim == 2:
        return np.array([(g.kty - W + iy) * g.iktx + ix for iy in range(2 * W) for ix in range(W)])
    return np.array([((g.ktz - W + iz) * g.ikty + g.ktz - W + iy) * g.iktx + ix
                     for iz in range(2 * W) for iy in range(2 * W) for ix in range(W)])


def shift_grid(nshift, alpha):
    nx, ny, nz = nshift
    sx = 2 * np.pi * np.arange(nx) / (nx * alpha)
    sy = 2 * np.pi * np.arange(ny) / ny
    sz = 2 * np.pi * np.arange(nz) / nz
    # flattened s = x + NX (y + NY z)
    SX = np.empty(nx * ny * nz)
    SY = np.empty_like(SX)
    SZ = np.empty_like(SX)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                s = x + nx * (y + ny * z)
                SX[s], SY[s], SZ[s] = sx[x], sy[y], sz[z]
    return SX, SY, SZ
