"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

The oracle's own tests (reconstruction, operator, limiter, diffusion, coupled
law, phase) run again in a child process against a build of mp_oracle.c with
-fsanitize=address,undefined and -fno-sanitize-recover: any out-of-bounds
access, use after free or undefined operation aborts the child and fails this
test.  Host code only (the device kernels are checked by their own tests).
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")
FILES = ["test_oracle_weno.py", "test_oracle_operator.py", "test_oracle_limiter.py",
         "test_oracle_diffusion.py", "test_oracle_coupled.py", "test_oracle_phase.py"]


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if out.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_tests_pass_under_asan_and_ubsan():
    asan = _runtime("libasan.so") if shutil.which("gcc") else None
    if asan is None or shutil.which("make") is None:
        pytest.skip("gcc's sanitizer runtime or make is not available")
    build = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ)
    env.update(MPO_ORACLE_LIB=os.path.join(ORACLE, "_san", "liboracle_mp.so"),
               LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               OMP_NUM_THREADS="4")
    # -s: a sanitizer report ends the child at once, so it must not sit in
    # pytest's own capture
    run = subprocess.run([sys.executable, "-m", "pytest", "-q", "-s", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                          *[os.path.join(HERE, f) for f in FILES]],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    text = run.stdout + run.stderr
    assert "runtime error" not in text and "AddressSanitizer" not in text, text[-4000:]
    assert run.returncode == 0, text[-4000:]
    assert " passed" in run.stdout
