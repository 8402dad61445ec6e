"""The CPU tests of the oracle and of the library's host C++ (scene
constants, scene files, drops and emitters) under AddressSanitizer and
UndefinedBehaviorSanitizer: `make -C oracle asan` builds both with
-fsanitize=address,undefined (host code only), and the tests run again in a
child process bound to those builds.  Any sanitizer report fails the run
(ASan aborts, UBSan halts on the first error); leak checking is off because
the Python interpreter itself is not built for it.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
FILES = ["tests/test_params_scene.py", "tests/test_scene_xml.py", "tests/test_inject.py",
         "tests/test_oracle_closed_forms.py", "tests/test_boundary_shapes.py", "tests/test_colours.py"]


def runtime(name: str) -> str:
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True, check=True)
    return out.stdout.strip()


def test_host_code_and_oracle_under_sanitizers():
    try:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "asan"], check=True, capture_output=True,
                       text=True)
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.fail(f"sanitizer build failed: {getattr(e, 'stderr', e)}")
    asan = runtime("libasan.so")
    env = dict(os.environ)
    env.update(
        LD_PRELOAD=asan,
        ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:allocator_may_return_null=1",
        UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
        PIBITI_ORACLE_LIB=str(ROOT / "oracle" / "_asan" / "liboracle.so"),
        PIBITI_SPH_HOST_LIB=str(ROOT / "oracle" / "_asan" / "libpibiti_host.so"),
        OMP_NUM_THREADS="2",
    )
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider", *FILES]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail
    assert " passed" in r.stdout, tail
