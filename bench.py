"""Benchmark: 3D Navier-Stokes 256^3 pseudospectral time steps per second on one MI355X.

Workload (one "step" = one Heun + Crank-Nicolson time step, two right-hand
sides): 256^3 grid, 2/3 spherical dealiasing (KT=85, 2,514,726 stored modes
per field), nu = 1/100, dt = 0.002, Kolmogorov forcing f = (sin 4y, 0, 0),
seeded random solenoidal initial velocity with E(k) ~ k^4 exp(-2(k/4)^2),
E = 0.5 (synthetic data).  The state is resident in HBM before timing starts.

`value` is the plain FP64 form's steps/s.  The FP32-bulk form (FFTs and
products in FP32, state/update/diagnostics FP64) is timed on the same input
and reported under "fp32_bulk" with its speed-up and its energy error
against the FP64 GPU run.  "roofline" is the dominant kernel's achieved HBM
bandwidth (algorithmic bytes per launch / mean launch time from HIP events on
the context stream during the timed steps); "model" gives the whole-step
fraction of the 8 TB/s roofline with the 5-pass pruned-pencil byte model.
"cpu_baseline" is the FP64 C oracle (test infrastructure) timed on this
host's cores on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--n N] [--no-cpu] [--no-fp32]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def model_bytes_step(n, b):
    """5-pass pruned-pencil bytes per Heun step (2 RHS) for an n^3 grid."""
    kt = n // 3
    iktx, ikty = kt + 1, 2 * kt + 1
    nkt = iktx * ikty * ikty
    Z, Y = iktx * ikty * n, iktx * n * n
    S, Fi, Fo = 3, 6, 3
    rhs = b * (S * nkt + 2 * Fi * Z + 2 * Fi * Y + 2 * Fo * Y + 2 * Fo * Z)
    stage = 6 * S * nkt * 16          # state/RH traffic stays FP64
    return 2 * rhs + stage


def time_form(P, st0, n, precision, steps, warmup, profile, tune=True):
    s = P.Stepper(flavour="ns3d", n=n, v2=0.01, dt=0.002, force="kolmogorov", precision=precision)
    s.set_state(st0)
    tuned = s.autotune(steps=4) if tune else None
    if tuned:
        log("bench: autotune %s:\n%s" % (precision, tuned.rstrip()))
    s.run(warmup)
    s.sync()
    if profile:
        s.profile(True)
    t0 = time.perf_counter()
    s.run(steps)
    s.sync()
    dt = time.perf_counter() - t0
    prof = s.profile_report() if profile else []
    e, z, _ = s.diag()
    variant = s.get_variant()
    s.close()
    return dt, prof, e, variant


def cpu_baseline(n, steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cores = O.lib().or_get_threads()
    st0 = O.ic_solenoidal(O.grid3(n, n, n), 20240601)
    orc = O.NS3D(n, v2=0.01, dt=0.002, force=3)
    orc.run(st0, 1)                      # warm caches / twiddles
    t0 = time.perf_counter()
    orc.run(st0, steps)
    dt = time.perf_counter() - t0
    return {"value": steps / dt, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": "%d Heun+CN steps of the %d^3 workload, FP64 C oracle, OpenMP %d threads" % (steps, n, cores)}


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
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import psgpu_amd as P
    n = a.n
    g = P.hgrid("ns3d", n, n, n)
    st0 = P.ic_solenoidal(g, 20240601, 0.5)
    log("bench: ns3d %d^3 fp64, warmup %d, timed %d steps" % (n, a.warmup, a.steps))
    t64, prof64, e64, var64 = time_form(P, st0, n, "fp64", a.steps, a.warmup, True, not a.no_tune)
    sps64 = a.steps / t64
    log("bench: fp64 %.2f steps/s" % sps64)
    fp32 = None
    if not a.no_fp32:
        log("bench: ns3d %d^3 fp32-bulk" % n)
        t32, prof32, e32, var32 = time_form(P, st0, n, "fp32", a.steps, a.warmup, True, not a.no_tune)
        sps32 = a.steps / t32
        fp32 = {"steps_per_s": sps32, "ms_per_step": 1e3 * t32 / a.steps, "speedup_vs_fp64": sps32 / sps64,
                "energy_rel_err_vs_fp64_gpu": abs(e32 - e64) / e64, "steps": a.steps, "xpass_variant": var32,
                "kernels": prof32}
        log("bench: fp32-bulk %.2f steps/s (x%.2f)" % (sps32, sps32 / sps64))
    # dominant kernel of the FP64 step
    dom = max(prof64, key=lambda r: r["avg_ms"] * r["count"]) if prof64 else None
    roof = None
    if dom:
        ach = dom["bytes"] / (dom["avg_ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom["kernel"], "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS, "traffic": None, "bytes_per_launch": dom["bytes"],
                "avg_launch_ms": dom["avg_ms"]}
    mb = model_bytes_step(n, 16)
    model = {"gb_per_step": mb / 1e9, "achieved_gbs": mb * sps64 / 1e9, "frac": mb * sps64 / 1e9 / PEAK_HBM_GBS}
    cpu = None
    if not a.no_cpu:
        log("bench: cpu baseline (%d steps)" % a.cpu_steps)
        cpu = cpu_baseline(n, a.cpu_steps)
    out = {"metric": "3D NS %d^3 timesteps/s (FP64 plain; FP32-bulk beside)" % n, "value": sps64,
           "unit": "steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * t64 / a.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "ns3d_heun_cn_%d3_kolmogorov" % n, "grid": [n, n, n], "nu": 0.01, "dt": 0.002,
                      "forcing": "sin(4y) x", "ic": "solenoidal E(k)~k^4exp(-2(k/4)^2), E=0.5, seed 20240601",
                      "members": 1, "xpass_variant": var64},
           "roofline": roof, "model": model, "fp32_bulk": fp32, "kernels_fp64": prof64, "cpu_baseline": cpu}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
