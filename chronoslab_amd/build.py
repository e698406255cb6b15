"""In-tree build of the native libraries.

    python -m chronoslab_amd.build        # both libraries, every time it is called

libchronos_amg.so   hipcc, gfx950, from csrc/gpu/*.hip
libchronos_setup.so g++ -fopenmp, from csrc/host/*.cpp
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CHRONOS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 ops stay off in every device compilation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no FMA contraction: every product is rounded before it is added, so the block and
    # single-vector kernels (and the FP64 oracle) round identically
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, file=sys.stderr)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_setup() -> str:
    out = os.path.join(HERE, "libchronos_setup.so")
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "host", "*.cpp")))
    _run(["g++", *HOST_FLAGS, f"-I{INC}", f"-I{os.path.join(HERE, 'csrc', 'common')}", "-o", out + ".tmp", *srcs])
    os.replace(out + ".tmp", out)
    return out


def build_amg() -> str:
    out = os.path.join(HERE, "libchronos_amg.so")
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "gpu", "*.hip")))
    _run([HIPCC, *HIP_FLAGS, f"-I{INC}", f"-I{os.path.join(HERE, 'csrc', 'gpu')}",
          f"-I{os.path.join(HERE, 'csrc', 'common')}", "-o", out + ".tmp", *srcs])
    os.replace(out + ".tmp", out)
    return out


def build_all() -> None:
    build_setup()
    build_amg()


if __name__ == "__main__":
    build_all()
