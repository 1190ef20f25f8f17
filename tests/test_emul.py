"""The product C-ABI built for the CPU (tests/emul: OpenMP workgroups, fiber
threads switching at barriers, AddressSanitizer) run over every flavour,
x-pass shape, graph path, tangent-linear/adjoint and Newton entry point at
16^3-64^2 sizes against the FP64 oracle, and the pencil FFT alone at every
length 16..1024, points per thread and LDS swizzle against a direct DFT.
It is the bounds and barrier check of the kernels on the host; it takes
several minutes, so it runs only with PSG_EMUL=1 in the environment."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emul")


@pytest.mark.skipif(os.environ.get("PSG_EMUL") != "1", reason="host bounds-check build: set PSG_EMUL=1")
def test_emulated_kernels_in_bounds_and_match_oracle():
    subprocess.check_call(["make", "-s", "-C", HERE, "-j8"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:detect_stack_use_after_return=0")
    r = subprocess.run([os.path.join(HERE, "_build", "emul_check"), "quick"], env=env, capture_output=True,
                       text=True, timeout=3000)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout


@pytest.mark.skipif(os.environ.get("PSG_EMUL") != "1", reason="host bounds-check build: set PSG_EMUL=1")
def test_emulated_pencil_fft_all_lengths_and_swizzles():
    subprocess.check_call(["make", "-s", "-C", HERE, "-j8"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:detect_stack_use_after_return=0")
    r = subprocess.run([os.path.join(HERE, "_build", "fft_check")], env=env, capture_output=True, text=True,
                       timeout=1200)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
