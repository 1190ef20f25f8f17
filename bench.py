"""Benchmark: 3D Navier-Stokes 256^3 pseudospectral time steps per second on one MI355X.

Workload (one "step" = one Heun + Crank-Nicolson time step, two right-hand
sides): 256^3 grid, 2/3 spherical dealiasing (KT=85, 2,514,726 stored modes
per field), nu = 1/100, dt = 0.002, Kolmogorov forcing f = (sin 4y, 0, 0),
seeded random solenoidal initial velocity with E(k) ~ k^4 exp(-2(k/4)^2),
E = 0.5 (synthetic data).  The state is resident in HBM before timing starts.

`value` is the plain FP64 form's steps/s (captured step graphs, kernel shapes
autotuned first).  The FP32-bulk form (FFTs and products in FP32,
state/update/diagnostics FP64) is timed on the same input and reported under
"fp32_bulk" with its speed-up and its energy error against the FP64 GPU run,
and, under "energy_1000", the same error every 100 steps of a 1000-step run.
"roofline" is the dominant kernel's achieved HBM bandwidth: algorithmic bytes
per launch / mean launch time from HIP events on the context stream over a
separate profiled run of 20 steps; its "traffic" is the kernel's measured HBM
bytes from the committed rocprofv3 summary.  "model" is the whole-step fraction of the
8 TB/s roofline under the 5-pass pruned-pencil byte model.  "cpu_baseline" is
the FP64 C oracle (test infrastructure) timed on this host's cores on a
bounded sample of the same workload.

Further cases on the same line: "adjoint_128" (gradient of ||Phi_T(u0)-u0||^2
at 128^3 over 200 steps with checkpoints every 20, error against fourth-order
central differences along 16 seeded solenoidal directions, FP64 and FP32-bulk),
"bous_ensemble_512" (3D Boussinesq 512^3, 4 members stepped together, FP64 and
FP32-bulk) and "newton_2d" (2D Kolmogorov 512^2 Newton-GMRES-hookstep from the
closest K recurrence guesses, K = 1 by default and 0 to skip: seconds per
converged orbit with finite-difference, tangent-linear and FP32 tangent-linear Jv).

    python bench.py [--steps K] [--warmup W] [--n N] [--no-cpu] [--no-fp32] [--no-tune]
                    [--no-adjoint] [--no-bous] [--no-energy] [--newton K]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "3D NS 256\u00b3 timesteps/s FP64 vs FP32-bulk, %HBM roofline, Newton s/orbit"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# rows of the traffic summary for the FP64 pass names of the step profile
TRAFFIC_ROW = {"zinv": "k_zinv<double, 256", "yinv": "k_yinv<double, 256", "xpass": "k_xpass<double, 256",
               "yfwd": "k_yfwd<double, 256", "zfwd_update": "k_zfwd<double, 256"}


def measured_traffic(kernel_prefix):
    """HBM bytes per launch of a kernel as measured by the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs summarised in profiles/ns3d_256_hbm_traffic.txt (calibrated
    MB read + written); None when the summary or the kernel's row is absent."""
    path = os.path.join(ROOT, "profiles", "ns3d_256_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(kernel_prefix):
            vals = line[60:].split()
            return (float(vals[0]) + float(vals[2])) * 1e6
    return None


def model_bytes_step(n, b, s=16):
    """5-pass pruned-pencil bytes per Heun step (2 RHS) for an n^3 grid: pass buffers of
    b bytes per complex value, state of s bytes (FP64 16; the FP32-bulk working state 8)."""
    kt = n // 3
    iktx, ikty = kt + 1, 2 * kt + 1
    nkt = iktx * ikty * ikty
    Z, Y = iktx * ikty * n, iktx * n * n
    S, Fi, Fo = 3, 6, 3
    # pass 1 reads the state; the pass buffers are in the pass precision b
    rhs = s * S * nkt + b * (2 * Fi * Z + 2 * Fi * Y + 2 * Fo * Y + 2 * Fo * Z)
    # Heun state/RH traffic: predictor reads w^n, writes w* and RH; corrector reads w*, RH, writes
    # w^{n+1}; RH is kept in the pass precision
    stage = S * nkt * (4 * s + 2 * b)
    return 2 * rhs + stage


def time_form(P, st0, n, precision, steps, warmup, tune=True, prof_steps=20):
    """Timed steps (captured step graphs, no per-kernel events), then a separate
    profiled run whose HIP events give each kernel's mean launch time."""
    s = P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, force="kolmogorov", precision=precision)
    s.set_state(st0)
    tuned = s.autotune(steps=4) if tune else None
    if tuned:
        log("bench: autotune %s:\n%s" % (precision, tuned.rstrip()))
    s.run(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.run(steps)
    s.sync()
    dt = time.perf_counter() - t0
    e, z, _ = s.diag()
    prof = []
    if prof_steps:
        s.profile(True)
        s.run(prof_steps)
        s.sync()
        prof = s.profile_report()
        s.profile(False)
    variant = {"xpass": s.get_variant(0), "zfwd": s.get_variant(1), "slab": s.get_variant(2)}
    s.close()
    return dt, prof, e, variant


def energy_drift(P, st0, n, nsteps=1000, every=100):
    """FP32-bulk against plain FP64 over nsteps Heun steps from the same state:
    relative energy difference every `every` steps (the target is <= 1e-6 at 1000)."""
    series = {}
    for precision in ("fp64", "fp32"):
        with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, force="kolmogorov", precision=precision) as s:
            s.set_state(st0)
            es = []
            for _ in range(nsteps // every):
                s.run(every)
                es.append(s.diag()[0])
            series[precision] = es
    rel = [abs(a - b) / abs(b) for a, b in zip(series["fp32"], series["fp64"])]
    return {"steps": nsteps, "every": every, "energy_fp64": series["fp64"], "rel_err": rel, "rel_err_final": rel[-1],
            "rel_err_max": max(rel)}


def adjoint_case(P, n=128, nsteps=200, ckpt=20, ndir=16, eps=1e-4, eps2=1e-6):
    """Gradient of J = ||Phi_T(u0) - u0||^2 at 128^3, 200 steps, checkpoints every 20,
    against fourth-order central differences along seeded solenoidal directions (FD
    with the FP64 stepper); FP64 and FP32-bulk adjoints.  With J ~ 10 and h = 1e-4 the
    FD's round-off is ~1e-10 and its truncation error O(h^4).  The configuration's
    literal check, second-order central differences at eps2 = 1e-6, is reported
    beside it ("cd2_*"): its round-off, ~5e-9, is 5e-6 relative on a direction whose
    derivative is 1e-3, so its error is also stated over ||grad||."""
    g = P.hgrid("ns3d", n, n, n)
    L = P.tables(g)[3].astype(float)
    z0 = P.ic_solenoidal(g, 20240603, 0.5)
    out = {"grid": [n, n, n], "steps": nsteps, "dt": 0.002, "ckpt": ckpt, "directions": ndir, "eps": eps,
           "fd_order": 4}
    with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, force="kolmogorov") as s:
        def J(x):
            s.set_state(x)
            s.run(nsteps)
            r = (s.get_state() - x) * L[None, :]
            return float(np.sum(np.abs(r) ** 2))
        s.residual_grad(z0, nsteps, ckpt)       # warm
        s.sync()
        t0 = time.perf_counter()
        J0, g64, _ = s.residual_grad(z0, nsteps, ckpt)
        out["fp64_grad_seconds"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        J(z0)
        out["forward_seconds"] = time.perf_counter() - t0
        errs64, errs32, fds, nerr64, cd2, cd2n = [], [], [], [], [], []
        gnorm = float(np.sqrt(np.sum(np.abs(g64) ** 2 * L[None, :])))
        with P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, force="kolmogorov", precision="fp32") as s32:
            s32.residual_grad(z0, nsteps, ckpt)
            t0 = time.perf_counter()
            _, g32, _ = s32.residual_grad(z0, nsteps, ckpt)
            out["fp32_grad_seconds"] = time.perf_counter() - t0
        for k in range(ndir):
            v = P.ic_solenoidal(g, 777000 + k, 0.5)
            v /= np.sqrt(np.sum(np.abs(v) ** 2))
            fd = (8 * (J(z0 + eps * v) - J(z0 - eps * v)) - (J(z0 + 2 * eps * v) - J(z0 - 2 * eps * v))) / (12 * eps)
            ip = lambda a: float(np.sum((np.conj(a) * v).real * L[None, :]))
            fds.append(fd)
            c2 = (J(z0 + eps2 * v) - J(z0 - eps2 * v)) / (2 * eps2)
            cd2.append(abs(ip(g64) - c2) / abs(c2))
            cd2n.append(abs(ip(g64) - c2) / gnorm)
            errs64.append(abs(ip(g64) - fd) / abs(fd))
            errs32.append(abs(ip(g32) - fd) / abs(fd))
            nerr64.append(abs(ip(g64) - fd) / gnorm)
    # per direction |<g,v> - fd| / |fd|; and the same error over ||g|| (|v| = 1), which does not
    # inflate when a direction happens to be nearly orthogonal to the gradient
    out.update({"J": J0, "grad_norm": gnorm, "fp64_max_rel_err_vs_fd": max(errs64),
                "fp64_median_rel_err_vs_fd": float(np.median(errs64)), "fp32_max_rel_err_vs_fd": max(errs32),
                "fp64_max_err_over_grad_norm": max(nerr64), "fd_values": fds,
                "cd2_eps": eps2, "cd2_fp64_max_rel_err": max(cd2), "cd2_fp64_median_rel_err": float(np.median(cd2)),
                "cd2_fp64_max_err_over_grad_norm": max(cd2n)})
    return out


def newton_case(P, n=512, t_spin=100.0, t_window=300.0, dt_sample=0.25, nguess=4, max_newton=20,
                tmin=5.0, tmax=30.0, save_orbit=None):
    """2D Kolmogorov flow n^2, Re=40, sin(4y) forcing, dt=0.005: spin up a seeded
    random-phase state, take the closest shift-minimised near-recurrences with
    T in [tmin, tmax] as guesses, and run Newton-GMRES-hookstep (Krylov 40,
    residual 1e-10, at most max_newton iterations) on all guesses with
    finite-difference Jv, then on the same guesses with tangent-linear Jv in FP64
    and in FP32-bulk passes (jv 0, 1, 2), one guess after the other on one
    context; s/orbit = total time / converged, and "paired" lists the guesses
    the finite difference and a tangent-linear form both converged, with times."""
    from psgpu_amd import recurrence as R
    g = P.hgrid("ns2d", n, n)
    L = P.tables(g)[3]
    z0 = P.ic_ns2d_cfg(g, 20240607, 1.0)
    out = {"grid": [n, n], "re": 40.0, "dt": 0.005, "t_spin": t_spin, "window": t_window, "krylov": 40,
           "netol": 1e-10, "max_newton": max_newton}
    kw = dict(flavour="ns2d", n=n, v2=1 / 40.0, dt=0.005, kf=4.0, ampfor=-2.0)
    with P.Stepper(**kw) as s:
        t0 = time.perf_counter()
        times, snaps = R.record(s, z0, t_spin, t_window, dt_sample)
        out["record_seconds"] = time.perf_counter() - t0
        out["dns_steps_per_s"] = (t_spin + t_window) / 0.005 / out["record_seconds"]
        t0 = time.perf_counter()
        hits = R.best_recurrences(times, snaps, g.iktx, L, tmin, tmax, top=nguess)
        out["search_seconds"] = time.perf_counter() - t0
        out["guesses"] = hits
        log("bench: newton guesses " + ", ".join("t=%.2f T=%.3f s=%.3f d=%.3f" % (h["t"], h["T"], h["shift"],
                                                                                  h["dist"]) for h in hits))
        runs = []
        names = {0: "finite-difference", 1: "tangent-linear", 2: "tangent-linear-fp32"}
        for jv in (0, 1, 2):
            name = names[jv]
            rs = []
            for k, h in enumerate(hits):
                z, res = s.newton(snaps[h["i"]][None], T=h["T"], sx=h["shift"], jv=jv, netol=1e-10, max_gmres=40,
                                  max_newton=max_newton, dt0=0.005, verbose=1)
                res["jv"] = name
                res["guess"] = k
                if save_orbit and jv == 0 and res["status_name"] == "CONVERGED" and not os.path.exists(save_orbit):
                    L1 = L == 1
                    np.savez_compressed(save_orbit, z_active=z[0][L1], T=res["T"], sx=res["sx"], ner=res["ner"],
                                        n=n, re=40.0, dt0=0.005)
                    log("bench: converged orbit saved to %s" % save_orbit)
                log("bench: newton guess %d (T=%.3f) %s: %s ner=%.2e T=%.6f in %.1f s, %d iterations" %
                    (k, h["T"], name, res["status_name"], res["ner"], res["T"], res["seconds"], res["newton_iters"]))
                rs.append(res)
            runs += rs
            conv = [r for r in rs if r["status_name"] == "CONVERGED"]
            total = sum(r["seconds"] for r in rs)
            out[name] = {"converged": len(conv), "tried": len(rs), "seconds_total": total,
                         "s_per_orbit": total / len(conv) if conv else None,
                         "newton_iters": [r["newton_iters"] for r in rs], "status": [r["status_name"] for r in rs]}
        out["runs"] = runs
        conv = lambda nm: {r["guess"]: r for r in runs if r["jv"] == nm and r["status_name"] == "CONVERGED"}
        fd = conv("finite-difference")
        out["paired"] = []
        for nm in ("tangent-linear", "tangent-linear-fp32"):
            tl = conv(nm)
            out["paired"] += [{"guess": k, "T": fd[k]["T"], "jv": nm, "fd_seconds": fd[k]["seconds"],
                               "tl_seconds": tl[k]["seconds"], "fd_iters": fd[k]["newton_iters"],
                               "tl_iters": tl[k]["newton_iters"], "fd_phi_evals": fd[k]["phi_evals"],
                               "tl_phi_evals": tl[k]["phi_evals"]} for k in sorted(set(fd) & set(tl))]
    return out


def newton_batch_case(P, n=512, nguess=8, lanes=8, max_newton=20, t_spin=100.0, t_window=300.0, dt_sample=0.25,
                      tmin=5.0, tmax=30.0, sequential=True):
    """The Newton case with many guesses at once (psg_newton_batch): the nguess closest
    recurrences of the same DNS as newton_case, solved with finite-difference Jv on one
    context of `lanes` members, and (sequential) one after the other on a batch-1 context."""
    from psgpu_amd import recurrence as R
    g = P.hgrid("ns2d", n, n)
    L = P.tables(g)[3]
    z0 = P.ic_ns2d_cfg(g, 20240607, 1.0)
    kw = dict(flavour="ns2d", n=n, v2=1 / 40.0, dt=0.005, kf=4.0, ampfor=-2.0)
    with P.Stepper(**kw) as s:
        times, snaps = R.record(s, z0, t_spin, t_window, dt_sample)
    hits = R.best_recurrences(times, snaps, g.iktx, L, tmin, tmax, top=nguess)
    zs = np.array([snaps[h["i"]][None] for h in hits])
    Ts = np.array([h["T"] for h in hits])
    ss = np.array([h["shift"] for h in hits])
    opts = dict(jv=0, netol=1e-10, max_gmres=40, max_newton=max_newton, dt0=0.005, verbose=1)
    out = {"grid": [n, n], "guesses": len(hits), "lanes": lanes, "T": list(Ts), "max_newton": max_newton,
           "T_range": [tmin, tmax]}
    log("bench: newton batch guesses T = " + ", ".join("%.3f" % t for t in Ts))
    with P.Stepper(batch=lanes, **kw) as sb:
        t0 = time.perf_counter()
        zb, Tb, sxb, rb = sb.newton_batch(zs, Ts, sx=ss, **opts)
        out["batched_seconds"] = time.perf_counter() - t0
    out["batched"] = [{"status": r["status_name"], "ner": r["ner"], "T": r["T"], "iters": r["newton_iters"],
                       "phi_evals": r["phi_evals"]} for r in rb]
    out["batched_converged"] = sum(r["status_name"] == "CONVERGED" for r in rb)
    log("bench: newton batch of %d guesses on %d lanes: %.1f s, %d converged" % (
        len(hits), lanes, out["batched_seconds"], out["batched_converged"]))
    if sequential:
        with P.Stepper(**kw) as s:
            t0 = time.perf_counter()
            rs = [s.newton(zs[k], T=Ts[k], sx=ss[k], **opts)[1] for k in range(len(hits))]
            out["sequential_seconds"] = time.perf_counter() - t0
        out["sequential"] = [{"status": r["status_name"], "ner": r["ner"], "T": r["T"], "iters": r["newton_iters"]}
                             for r in rs]
        out["sequential_converged"] = sum(r["status_name"] == "CONVERGED" for r in rs)
        log("bench: the same guesses one after the other: %.1f s, %d converged" % (
            out["sequential_seconds"], out["sequential_converged"]))
    return out


def bous_ensemble(P, n=512, members=4, steps=20, warmup=2, tune=True):
    """3D Boussinesq n^3, N=Ri=1, Pr=1, nu=kappa=1/400, Kolmogorov forcing, RK4 with
    per-stage CN, dt=0.002, theta=90 (radians, as the reference passes it): `members`
    seeded ensemble members stepped together, captured step graphs, FP64 and FP32-bulk,
    x-pass shape autotuned per precision; the energy comparison is after warmup + steps."""
    g = P.hgrid("bous", n, n, n)
    out = {"grid": [n, n, n], "members": members, "steps": steps, "dt": 0.002, "nu": 1 / 400.0, "ri": 1.0,
           "sc": 1.0, "theta": 90.0}
    sts = [P.ic_solenoidal(g, 20240605 + m, 0.5, 0.1) for m in range(members)]
    for prec in ("fp64", "fp32"):
        with P.Stepper(flavour="bous", n=n, v2=1 / 400.0, sc=1.0, ri=1.0, theta=90.0, dt=0.002, batch=members,
                       force="kolmogorov", precision=prec) as s:
            for m in range(members):
                s.set_state(sts[m], m)
            var = 0
            if tune:
                rep = s.autotune(2)
                log("bench: bous %s autotune: %s" % (prec, rep.strip().replace("\n", "; ")))
                var = s.get_variant(0)
                for m in range(members):
                    s.set_state(sts[m], m)
            s.set_graph(2)
            s.run(warmup)
            s.sync()
            t0 = time.perf_counter()
            s.run(steps)
            s.sync()
            dt = time.perf_counter() - t0
            e = [s.diag(m)[0] for m in range(members)]
        out[prec] = {"ensemble_steps_per_s": steps / dt, "member_steps_per_s": members * steps / dt,
                     "ms_per_step": 1e3 * dt / steps, "energy": e, "xpass_variant": var}
        log("bench: bous %d^3 x%d %s %.2f ensemble steps/s (x-pass variant %d)" % (n, members, prec, steps / dt,
                                                                                   var))
    out["fp32_speedup"] = out["fp32"]["ensemble_steps_per_s"] / out["fp64"]["ensemble_steps_per_s"]
    out["fp32_energy_rel_err_vs_fp64"] = max(abs(a - b) / b for a, b in zip(out["fp32"]["energy"],
                                                                             out["fp64"]["energy"]))
    return out


def recurrence_case(P, n=256, stores=500, cpu_sample=12, reps=5):
    """Recurrence search at 256^3 with the reference's sizes (500 history entries,
    the 2048-mode subset, 30x2x2 shifts): GPU time of a NOUT block (a check against
    the full ring, 499 stores, and the push) averaged over `reps` blocks, and as its
    CPU baseline the oracle's sequential check on one host core over a bounded
    sample of stores, scaled to 499."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    st = P.Stepper(flavour="ns3d", n=n, dt=0.002, v2=0.01)
    g = st.grid
    rng = np.random.default_rng(5)
    z = (rng.standard_normal((3, st.nkt)) + 1j * rng.standard_normal((3, st.nkt))) * st.mask[None]
    st.set_state(z)
    rec = P.Recurrence(st, nout=10, store_size=stores, keep_states=0, threshold=0.0)
    ph = np.exp(1j * rng.standard_normal(st.nkt) * 0.01)
    for e in range(1, stores + 1):
        z = z * ph[None]
        st.set_state(z)
        rec.update(10 * e, 0.002)
    t0 = time.perf_counter()
    for e in range(stores + 1, stores + 1 + reps):
        _, _, c = rec.update(10 * e, 0.002)
    gpu_ms = 1e3 * (time.perf_counter() - t0) / reps
    rec.close()
    st.close()
    # oracle: one check over cpu_sample + 1 entries (a ring of that size), one thread
    og = O.grid3(n, n, n)
    kx, ky, kz, L = O.tables(og)
    O.lib().or_set_threads(1)
    small = (rng.standard_normal((3, og.nkt)) + 1j * rng.standard_normal((3, og.nkt))) * L
    orr = O.Rec(0, og, L, small, store_size=cpu_sample + 1, nout=10, dt=0.002, threshold=0.0)
    for e in range(1, cpu_sample + 1):
        orr.update(10 * e, 0.002, small)
    t0 = time.perf_counter()
    orr.update(10 * (cpu_sample + 1), 0.002, small)
    cpu_s = time.perf_counter() - t0
    del orr
    O.lib().or_set_threads(os.cpu_count() or 1)
    per_store = cpu_s / (cpu_sample - 1)
    return {"grid": [n, n, n], "stores": c["compared"], "shifts": 120, "subset_modes": 2048,
            "gpu_ms_per_check": gpu_ms, "cpu_oracle_ms_per_check_1core": 1e3 * per_store * (stores - 1),
            "cpu_sample_stores": cpu_sample - 1}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(n, steps, steps_1core=1):
    """The FP64 C oracle on this host's cores (OpenMP, all threads it is given), and
    the same on one thread, over a bounded sample of the workload's steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cores = O.lib().or_get_threads()
    st0 = O.ic_solenoidal(O.grid3(n, n, n), 20240601)
    orc = O.NS3D(n, v2=0.01, dt=0.002, force=3)
    orc.run(st0, 1)                      # warm caches / twiddles
    t0 = time.perf_counter()
    orc.run(st0, steps)
    dt = time.perf_counter() - t0
    O.lib().or_set_threads(1)
    t0 = time.perf_counter()
    orc.run(st0, steps_1core)
    dt1 = time.perf_counter() - t0
    O.lib().or_set_threads(cores)
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"value": steps / dt, "unit": "steps/s", "cores": cores, "kind": "port",
            "value_1core": steps_1core / dt1, "cpu_model": cpu_model(), "nproc": nproc,
            "cpu_count": os.cpu_count(), "fft": "project-written FP64 FFT (no FFTW on the host)",
            "sample": "%d Heun+CN steps of the %d^3 workload, FP64 C oracle, OpenMP %d threads; %d step(s) on "
                      "1 thread" % (steps, n, cores, steps_1core)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-tune", action="store_true", help="keep the default kernel shapes (no autotune)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 128^3 adjoint-gradient case")
    ap.add_argument("--no-bous", action="store_true", help="skip the 512^3 Boussinesq ensemble case")
    ap.add_argument("--bous-steps", type=int, default=20,
                    help="timed Boussinesq ensemble steps (the configuration's run is 500)")
    ap.add_argument("--no-rec", action="store_true", help="skip the 256^3 recurrence-search timing")
    ap.add_argument("--no-energy", action="store_true",
                    help="skip the 1000-step FP32-bulk energy series (profiled runs)")
    ap.add_argument("--newton", type=int, default=1, metavar="NGUESS",
                    help="2D 512^2 Newton case (FD vs TL Jv) on NGUESS recurrence guesses (0: skip)")
    ap.add_argument("--newton-n", type=int, default=512)
    ap.add_argument("--newton-batch", type=int, default=0, metavar="LANES",
                    help="also solve --newton-guesses guesses together on LANES members (0: skip)")
    ap.add_argument("--newton-guesses", type=int, default=8)
    ap.add_argument("--newton-batch-iters", type=int, default=10, help="Newton iterations per batched guess")
    ap.add_argument("--newton-batch-tmax", type=float, default=8.0, help="longest guess period of the batch")
    ap.add_argument("--newton-batch-no-seq", action="store_true",
                    help="skip the one-guess-at-a-time comparison of the batched case")
    ap.add_argument("--save-orbit", default=None, metavar="NPZ",
                    help="write the first orbit the finite-difference Newton converges (active modes, T, shift)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import psgpu_amd as P
    n = a.n
    g = P.hgrid("ns3d", n, n, n)
    st0 = P.ic_solenoidal(g, 20240601, 0.5)
    log("bench: ns3d %d^3 fp64, warmup %d, timed %d steps" % (n, a.warmup, a.steps))
    t64, prof64, e64, var64 = time_form(P, st0, n, "fp64", a.steps, a.warmup, not a.no_tune)
    sps64 = a.steps / t64
    log("bench: fp64 %.2f steps/s" % sps64)
    fp32 = None
    if not a.no_fp32:
        log("bench: ns3d %d^3 fp32-bulk" % n)
        t32, prof32, e32, var32 = time_form(P, st0, n, "fp32", a.steps, a.warmup, not a.no_tune)
        sps32 = a.steps / t32
        fp32 = {"steps_per_s": sps32, "ms_per_step": 1e3 * t32 / a.steps, "speedup_vs_fp64": sps32 / sps64,
                "energy_rel_err_vs_fp64_gpu": abs(e32 - e64) / e64, "steps": a.steps, "pass_variants": var32,
                "kernels": prof32}
        log("bench: fp32-bulk %.2f steps/s (x%.2f)" % (sps32, sps32 / sps64))
        if not a.no_energy:
            fp32["energy_1000"] = energy_drift(P, st0, n)
            log("bench: fp32-bulk energy vs fp64 after 1000 steps: %.2e (max over the run %.2e)" %
                (fp32["energy_1000"]["rel_err_final"], fp32["energy_1000"]["rel_err_max"]))
    # dominant kernel of the FP64 step
    dom = max(prof64, key=lambda r: r["avg_ms"] * r["count"]) if prof64 else None
    roof = None
    if dom:
        ach = dom["bytes"] / (dom["avg_ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom["kernel"], "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS,
                "traffic": measured_traffic(TRAFFIC_ROW.get(dom["kernel"], "-")) if n == 256 else None,
                "traffic_source": "profiles/ns3d_256_hbm_traffic.txt (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE)",
                "bytes_per_launch": dom["bytes"],
                "avg_launch_ms": dom["avg_ms"]}
    mb = model_bytes_step(n, 16)
    model = {"gb_per_step": mb / 1e9, "achieved_gbs": mb * sps64 / 1e9, "frac": mb * sps64 / 1e9 / PEAK_HBM_GBS}
    if fp32:
        mb32 = model_bytes_step(n, 8, 8)
        fp32["model"] = {"gb_per_step": mb32 / 1e9, "achieved_gbs": mb32 * fp32["steps_per_s"] / 1e9,
                         "frac": mb32 * fp32["steps_per_s"] / 1e9 / PEAK_HBM_GBS}
    adj = None
    if not a.no_adjoint:
        log("bench: adjoint gradient 128^3, 200 steps")
        adj = adjoint_case(P)
        log("bench: adjoint fp64 %.3f s, max rel err vs FD %.2e" % (adj["fp64_grad_seconds"],
                                                                     adj["fp64_max_rel_err_vs_fd"]))
    bous = None
    if not a.no_bous:
        log("bench: boussinesq 512^3 ensemble")
        bous = bous_ensemble(P, steps=a.bous_steps, tune=not a.no_tune)
    recur = None
    if not a.no_rec and n == 256:
        log("bench: recurrence search 256^3, 500-entry history")
        recur = recurrence_case(P, n)
        log("bench: recurrence check %.2f ms on the GPU (%d stores); oracle %.0f ms on one core" %
            (recur["gpu_ms_per_check"], recur["stores"], recur["cpu_oracle_ms_per_check_1core"]))
    newton = None
    if a.newton:
        log("bench: newton 2D %d^2, %d guesses" % (a.newton_n, a.newton))
        newton = newton_case(P, n=a.newton_n, nguess=a.newton, save_orbit=a.save_orbit)
    nbatch = None
    if a.newton_batch:
        log("bench: newton 2D %d^2, %d guesses on %d lanes" % (a.newton_n, a.newton_guesses, a.newton_batch))
        nbatch = newton_batch_case(P, n=a.newton_n, nguess=a.newton_guesses, lanes=a.newton_batch,
                                   max_newton=a.newton_batch_iters, tmax=a.newton_batch_tmax,
                                   sequential=not a.newton_batch_no_seq)
    cpu = None
    if not a.no_cpu:
        log("bench: cpu baseline (%d steps)" % a.cpu_steps)
        cpu = cpu_baseline(n, a.cpu_steps)
    out = {"metric": METRIC, "value": sps64,
           "unit": "steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * t64 / a.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "ns3d_heun_cn_%d3_kolmogorov" % n, "grid": [n, n, n], "nu": 0.01, "dt": 0.002,
                      "forcing": "sin(4y) x", "ic": "solenoidal E(k)~k^4exp(-2(k/4)^2), E=0.5, seed 20240601",
                      "members": 1, "pass_variants": var64},
           "roofline": roof, "model": model, "fp32_bulk": fp32, "kernels_fp64": prof64,
           "adjoint_128": adj, "bous_ensemble_512": bous, "recurrence_256": recur, "newton_2d": newton,
           "newton_2d_batch": nbatch,
           "cpu_baseline": cpu}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
