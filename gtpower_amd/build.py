"""Build the native parts in-tree.

  libgtpower_hip.so    the engine: C-ABI host runtime + gfx950 HIP kernels (hipcc)
  libgtpower_synth.so  synthetic design/trace generator (g++), test/bench input
  oracle/liboracle.so  the FP64 CPU restatement of the reference path (g++), built
                       by oracle/build.py; only tests, smoke() and bench.py's CPU
                       baseline use it

Usage: python -m gtpower_amd.build [--engine|--synth|--oracle]
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GTPOWER_ARCH", "gfx950")

ENGINE_SOURCES = ["gtp_api.cpp", "gtp_kernels.hip", "gtp_mlp.hip", "gtp_vcd.hip"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # packed fp32 ops are disabled for every HIP object of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-fvisibility=hidden", "-DGTP_BUILD_DLL",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_synth():
    out = PKG / "libgtpower_synth.so"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
          PKG / "synth" / "synth.cpp", "-o", out])
    return out


def build_engine():
    out = PKG / "libgtpower_hip.so"
    tmp = out.with_suffix(".so.tmp")
    srcs = [CSRC / s for s in ENGINE_SOURCES]
    _run([HIPCC, *HIP_FLAGS, f"-I{REPO / 'include'}", f"-I{CSRC}", *srcs, "-o", tmp, "-pthread"])
    os.replace(tmp, out)
    return out


def build_oracle():
    sys.path.insert(0, str(REPO))
    from oracle import build as ob  # noqa: E402
    return ob.build()


def main(argv):
    which = set(a.lstrip("-") for a in argv[1:]) or {"synth", "engine", "oracle"}
    if "synth" in which:
        build_synth()
    if "oracle" in which:
        build_oracle()
    if "engine" in which:
        build_engine()


if __name__ == "__main__":
    main(sys.argv)
