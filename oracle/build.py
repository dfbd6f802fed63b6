"""Build recipe of the oracle (test infrastructure): oracle/liboracle.so from
oracle/oracle.cpp (power path) and oracle/vcd.cpp (VCD reading) with g++, and
oracle/liboracle_f32.so, the power path built with FP32 values (a diagnostic).  Nothing of the reference is compiled: the
reference needs nvcc, Tcl, SWIG, flex/bison, CUDD and Eigen, none of which this
image has, so oracle/_ref stays empty."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build():
    out = HERE / "liboracle.so"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
           str(HERE / "oracle.cpp"), str(HERE / "vcd.cpp"), "-o", str(out)]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    # the FP32 diagnostic build of the power path (energy terms, per-gate sums
    # and per-cycle waveforms in float, as the reference computes them)
    f32 = HERE / "liboracle_f32.so"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-DORACLE_REAL=float",
           str(HERE / "oracle.cpp"), "-o", str(f32)]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    build()
