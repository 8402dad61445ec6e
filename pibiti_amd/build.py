"""Build the HIP step library (gfx950) in-tree.

`python -m pibiti_amd.build` compiles every HIP/C++ source of the step into
``pibiti_amd/libpibiti_sph.so`` with hipcc.  Floating-point flags are part of
the contract: no FMA contraction and correctly rounded fp32 division and
square root, so one thread per particle reproduces the CPU restatement of the
reference arithmetic bit for bit.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libpibiti_sph.so"
RUN_BIN = PKG / "_build" / "sph_run"  # C++ host driver (pibiti_amd/host/sph_run.cpp)

HIP_SOURCES = ["sph_kernels.hip", "sph_fast.hip", "sph_sort.hip", "sph_resort.hip", "sph_skin.hip", "sph_capi.hip",
               "sph_mlp.hip", "sph_learn.hip"]
CPP_SOURCES = ["sph_scene.cpp", "sph_xml.cpp", "sph_inject.cpp"]

# -packed-fp32-ops keeps the compiler from emitting packed fp32 ops.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core results in ordinary VGPRs where they fit (no AGPR copies
    # between the products and the vector code that consumes them)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def _run(cmd: list[str]) -> None:
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> Path:
    """Compile the library; returns the path of the shared object."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    inc = ["-I", str(ROOT / "include")]
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        _run([hipcc(), *HIP_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    for src in CPP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        # host-only code: plain C++ (no device pass), same fp rules
        _run([hipcc(), "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", *inc, "-x", "c++", "-c",
              str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    # the C++ host driver: plain host code over the C ABI
    _run(["g++", "-O2", "-std=c++17", *inc, str(PKG / "host" / "sph_run.cpp"), "-L", str(PKG),
          "-lpibiti_sph", "-Wl,-rpath,$ORIGIN/..", "-o", str(RUN_BIN)])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
