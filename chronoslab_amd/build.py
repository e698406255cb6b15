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


def source_digest(paths: list[str]) -> str:
    """SHA-256 over the named files' contents (in the given order)."""
    import hashlib

    h = hashlib.sha256()
    for p in paths:
        h.update(os.path.basename(p).encode())
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def setup_sources() -> list[str]:
    return sorted(glob.glob(os.path.join(HERE, "csrc", "host", "*.cpp")) +
                  glob.glob(os.path.join(HERE, "csrc", "host", "*.hpp")) +
                  glob.glob(os.path.join(HERE, "csrc", "common", "*.hpp")) +
                  [os.path.join(INC, "chronos_setup.h")])


def amg_sources() -> list[str]:
    return sorted(glob.glob(os.path.join(HERE, "csrc", "gpu", "*.hip")) +
                  glob.glob(os.path.join(HERE, "csrc", "gpu", "*.hpp")) +
                  glob.glob(os.path.join(HERE, "csrc", "common", "*.hpp")) +
                  [os.path.join(INC, "chronos_amg.h")])


def is_current(lib: str, sources: list[str]) -> bool:
    """True when lib exists and was built from the sources as they are now (the digest
    written beside it by the build matches)."""
    try:
        with open(lib + ".srchash") as f:
            return os.path.exists(lib) and f.read().strip() == source_digest(sources)
    except OSError:
        return False


def _stamp(lib: str, sources: list[str]) -> None:
    with open(lib + ".srchash", "w") as f:
        f.write(source_digest(sources) + "\n")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, file=sys.stderr)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


# AddressSanitizer + UndefinedBehaviorSanitizer build of the host library (CPU checking only;
# load it with libasan preloaded, e.g. LD_PRELOAD=$(g++ -print-file-name=libasan.so)).
SANITIZE_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                  "-fno-sanitize-recover=undefined"]


def build_setup(sanitize: bool = False) -> str:
    out = os.path.join(HERE, "libchronos_setup.so")
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "host", "*.cpp")))
    flags = [f for f in HOST_FLAGS if f != "-O3"] + SANITIZE_FLAGS if sanitize else HOST_FLAGS
    _run(["g++", *flags, f"-I{INC}", f"-I{os.path.join(HERE, 'csrc', 'common')}", "-o", out + ".tmp", *srcs])
    os.replace(out + ".tmp", out)
    _stamp(out, setup_sources())
    return out


def build_amg() -> str:
    """Each .hip source is compiled to an object of its own (in parallel), then linked."""
    from concurrent.futures import ThreadPoolExecutor

    out = os.path.join(HERE, "libchronos_amg.so")
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "gpu", "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    incs = [f"-I{INC}", f"-I{os.path.join(HERE, 'csrc', 'gpu')}", f"-I{os.path.join(HERE, 'csrc', 'common')}"]
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([HIPCC, *cflags, *incs, "-c", "-o", so[1], so[0]]), zip(srcs, objs)))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)
    _stamp(out, amg_sources())
    return out


def build_all() -> None:
    build_setup()
    build_amg()


if __name__ == "__main__":
    if "--sanitize-host" in sys.argv:
        build_setup(sanitize=True)
    else:
        build_all()
