"""Build libgoofy_hip.so (gfx950) in-tree, and the test-only oracle libraries.

Usage: python goofy_amd/build.py [--no-oracle]   (run as a script: importing the
package would load the library being rebuilt)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "goofy_amd"
CSRC = PKG / "csrc"
LIB = PKG / "libgoofy_hip.so"
ARCH = os.environ.get("GOOFY_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "capi.cpp", "codec_api.cpp", "batch.cpp", "serve.cpp", "codec_extra.hip", "probe.hip",
           "containers.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-fvisibility=hidden", "-pthread",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = True) -> Path:
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}", "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    tmp.replace(LIB)
    return LIB


def build_examples(verbose: bool = True) -> None:
    """tests/native/libdropin_example.so: a C++ caller of the drop-in API."""
    out = ROOT / "tests" / "native" / "libdropin_example.so"
    cmd = ["g++", "-std=c++17", "-O1", "-fPIC", "-shared", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "native" / "dropin_example.cpp"), f"-L{PKG}", "-lgoofy_hip",
           f"-Wl,-rpath,{PKG}", "-o", str(out)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_oracle(verbose: bool = True) -> None:
    cmd = ["make", "-C", str(ROOT / "oracle"), "-s", "all"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def main(argv: list[str]) -> int:
    build_library()
    build_examples()
    if "--no-oracle" not in argv:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
