"""Build the psgpu-mi355x shared libraries in-tree.

    python -m psgpu_amd.build [-v] [-j N]

Outputs (git-ignored, shipped with the tree):
    psgpu_amd/_lib/libpsg.so           namespaced C-ABI (include/psg.h, psg_host.h)
    psgpu_amd/_lib/libpsg_legacy_*.so  one per reference flavour, exporting the
                                       reference symbols (include/psg_legacy.h)
    psgpu_amd/_lib/psg_gmres           the Newton-GMRES driver (C restatement of the
                                       reference's src_gmres/main.F90 callers)
    psgpu_amd/_lib/psg_dns             the DNS driver (3D NS, Euler, Boussinesq, 2D NS main.F90)
HIP sources are compiled for gfx950 only; the C host layer with gcc.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-I" + os.path.join(ROOT, "include")]

CORE_HIP = ["psg_api.hip", "psg_ns3d.hip", "psg_ns2d.hip", "psg_bous.hip", "psg_newton.hip", "psg_tladj.hip",
            "psg_rec.hip", "psg_stats.hip"]
CORE_C = ["psg_host.c"]
CXX20_HIP = {"psg_newton.hip"}   # the Newton solver is a C++20 coroutine
# legacy shims: (library suffix, source)
LEGACY = [("ns3d_dns", "legacy/legacy_ns3d_dns.c"), ("ns3d_gmres", "legacy/legacy_ns3d_gmres.c"),
          ("euler", "legacy/legacy_euler.c"), ("ns2d_gmres", "legacy/legacy_ns2d_gmres.c"),
          ("ns2d_dns", "legacy/legacy_ns2d_dns.c"), ("bous_dns", "legacy/legacy_bous_dns.c"),
          ("bous_gmres", "legacy/legacy_bous_gmres.c")]


# C drivers restating the reference's Fortran callers: (executable, source, libraries)
DRIVERS = [("psg_gmres", "drivers/psg_gmres.c", ["psg"]),
           ("psg_dns", "drivers/psg_dns.c", ["psg", "psg_legacy_ns3d_dns", "psg_legacy_euler", "psg_legacy_ns2d_dns",
                                             "psg_legacy_bous_dns"])]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + out)
    if verbose and out.strip():
        print(out)
    return out


def _compile(src, verbose):
    base = os.path.splitext(os.path.basename(src))[0]
    obj = os.path.join(OBJ, base + ".o")
    if src.endswith(".hip"):
        flags = HIP_FLAGS
        if src in CXX20_HIP:   # coroutines
            flags = [f if f != "-std=c++17" else "-std=c++20" for f in HIP_FLAGS]
        cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = ["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd, verbose)
    return obj


def build(verbose=False, jobs=None):
    """Compile every HIP/C source and link the shared libraries (always rebuilds)."""
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = CORE_HIP + CORE_C + [s for _, s in LEGACY]
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = dict(zip(srcs, ex.map(lambda s: _compile(s, verbose), srcs)))
    core = [objs[s] for s in CORE_HIP + CORE_C]
    lib = os.path.join(OUT, "libpsg.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + core + ["-lm"], verbose)
    os.replace(tmp, lib)
    for suffix, src in LEGACY:
        out = os.path.join(OUT, "libpsg_legacy_%s.so" % suffix)
        _run(["gcc", "-shared", "-fPIC", "-o", out + ".tmp", objs[src], "-L" + OUT, "-lpsg",
              "-Wl,-rpath,$ORIGIN", "-lm"], verbose)
        os.replace(out + ".tmp", out)
    for name, src, libs in DRIVERS:
        out = os.path.join(OUT, name)
        _run(["gcc"] + C_FLAGS + [os.path.join(CSRC, src), "-o", out + ".tmp", "-L" + OUT]
             + ["-l" + l for l in libs] + ["-Wl,-rpath,$ORIGIN", "-lm"], verbose)
        os.replace(out + ".tmp", out)
    return lib


def clean():
    for d in (OUT, OBJ):
        shutil.rmtree(d, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-v", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    print(build(a.v, a.j))
    sys.exit(0)
