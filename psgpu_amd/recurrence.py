"""Near-recurrence helpers for Newton guesses.

The reference's own search (the shift sweep over a 500-entry DNS history,
its turning-point and output-sequence rules, UPO_ZK.out records;
3D_NavierStokes/src/functions.h:213-389, 3D_Boussinesq/src/functions.h:415-619,
2D_NavierStokes/src_RK4/functions.h:105-309) runs on the GPU as
psgpu_amd.Recurrence (libpsg's psg_rec_*).

This module adds a continuous-shift variant for 2D Kolmogorov flow, used by
the bench to pick its Newton guess from snapshots of the GPU stepper: the
shift that best aligns two states is found for all shifts at once, because
the shifted inner product is a Fourier series in s over the kx index,

    sum_L Re(conj(a) e^{-i kx s} b) = Re sum_kx e^{-i kx s} c(kx),
    c(kx) = sum_ky conj(a(kx,ky)) b(kx,ky) L(kx,ky),

evaluated on M equally spaced shifts by one FFT of length M per pair (the
reference tries NSHIFTX = 30 shifts per pair).  The shift convention
S_s z = exp(-i kx s) z is the Newton solver's (include/psg_newton.h), so a hit
(t, T, s) is a guess for S_s Phi_T(z) = z.  It also reads and writes
UPO_ZK.out stream records.
"""
import numpy as np


def record(stepper, z0, t_spin, t_window, dt_sample):
    """Spin a 2D stepper up from z0 for t_spin, then keep a snapshot every
    dt_sample over t_window.  Returns (times, snapshots[n, nkt])."""
    dt = stepper.cfg.dt
    stepper.set_state(np.asarray(z0).reshape(1, -1))
    nspin = int(round(t_spin / dt))
    if nspin:
        stepper.run(nspin)
    every = max(1, int(round(dt_sample / dt)))
    nsnap = int(round(t_window / dt_sample)) + 1
    snaps = np.empty((nsnap, stepper.nkt), np.complex128)
    times = np.empty(nsnap)
    for i in range(nsnap):
        if i:
            stepper.run(every)
        snaps[i] = stepper.get_state()[0]
        times[i] = t_spin + i * every * dt
    return times, snaps


def shifted_overlap(a, B, iktx, L, M=512):
    """max over M shifts s_j = 2 pi j / M of sum_L Re(conj(a) S_s b) for each row b of B;
    returns (overlap[nb], shift[nb])."""
    ikty = a.size // iktx
    w = (np.conj(a) * L).reshape(ikty, iktx)
    c = np.einsum("yx,nyx->nx", w, B.reshape(len(B), ikty, iktx))
    f = np.fft.fft(c, n=M, axis=1).real          # sum_kx c(kx) e^{-2 pi i kx j / M}
    j = np.argmax(f, axis=1)
    return f[np.arange(len(B)), j], 2 * np.pi * j / M


def best_recurrences(times, snaps, iktx, L, tmin, tmax, top=8, M=512, separation=None):
    """Pairs (t_i, T = t_j - t_i in [tmin, tmax]) ordered by the shift-minimised
    relative distance ||S_s z_j - z_i|| / ||z_i||.  For each t_i only interior
    local minima of the distance over T count (a minimum at the end of the T
    window reflects slow drift, not a return); at most one hit per `separation`
    time units of t_i (default tmin) so that the guesses are distinct.
    Returns a list of dicts {i, t, T, shift, dist}."""
    L = np.asarray(L, float)
    norms = np.sqrt(np.sum(np.abs(snaps) ** 2 * L[None, :], axis=1))
    hits = []
    for i in range(len(times)):
        js = np.nonzero((times - times[i] >= tmin) & (times - times[i] <= tmax))[0]
        if js.size < 3:
            continue
        ov, sh = shifted_overlap(snaps[i], snaps[js], iktx, L, M)
        d2 = norms[i] ** 2 + norms[js] ** 2 - 2 * ov
        interior = np.nonzero((d2[1:-1] < d2[:-2]) & (d2[1:-1] <= d2[2:]))[0] + 1
        if interior.size == 0:
            continue
        k = int(interior[np.argmin(d2[interior])])
        hits.append({"i": i, "t": float(times[i]), "T": float(times[js[k]] - times[i]), "shift": float(sh[k]),
                     "dist": float(np.sqrt(max(d2[k], 0.0)) / norms[i])})
    hits.sort(key=lambda h: h["dist"])
    sep = tmin if separation is None else separation
    out = []
    for h in hits:
        if all(abs(h["t"] - o["t"]) >= sep for o in out):
            out.append(h)
        if len(out) == top:
            break
    return out


# UPO_ZK.out: the reference's guess file, a C stream of records, each 7 doubles
# followed by the state (2D_NavierStokes/src_RK4/functions.h:294-301:
# time, v2, period, shift x, shift y, residual, time under threshold, Z(nkt);
# 3D_NavierStokes/src/functions.h:368-377: time, v2, period, shift x, y, z,
# residual, XII, ETA, ZET(nkt each)).  Complex values are (re, im) doubles.
UPOZK_SCALARS = 7


def write_upozk(path, hits, snaps, v2, append=False):
    """Write 2D recurrence hits (best_recurrences) as UPO_ZK.out records: the
    state at the start of the near-return, its time, the period and x shift,
    y shift 0, the relative distance as the residual and 0 for the time under
    threshold (the search has no threshold)."""
    with open(path, "ab" if append else "wb") as f:
        for h in hits:
            sc = np.array([h["t"], v2, h["T"], h["shift"], 0.0, h["dist"], 0.0], np.float64)
            f.write(sc.tobytes())
            f.write(np.ascontiguousarray(snaps[h["i"]], np.complex128).tobytes())


def read_upozk(path, nkt, nfields=1):
    """Records of a UPO_ZK.out file with nfields states of nkt modes each (1 in 2D,
    3 in 3D): a list of (scalars[7], state[nfields, nkt] complex)."""
    rec = UPOZK_SCALARS + 2 * nfields * nkt
    raw = np.fromfile(path, np.float64)
    if raw.size % rec:
        raise ValueError("%s: %d doubles is not a whole number of %d-double records" % (path, raw.size, rec))
    out = []
    for r in raw.reshape(-1, rec):
        out.append((r[:UPOZK_SCALARS].copy(), r[UPOZK_SCALARS:].view(np.complex128).reshape(nfields, nkt)))
    return out
