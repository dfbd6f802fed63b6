"""Host sanitizer builds of the C++ test infrastructure (CPU only).

The oracle (oracle/oracle.cpp, oracle/vcd.cpp, its own worker threads) and
the synthetic generator (gtpower_amd/synth/synth.cpp, its parallel_for) are
built with AddressSanitizer + UndefinedBehaviorSanitizer ("asan") or with
ThreadSanitizer ("tsan") into oracle/_san/, and a small workload -- a random
netlist and trace generated, evaluated by the oracle on several threads, and
a VCD text written and read back -- runs against them in a child Python
whose sanitizer runtime is preloaded.  run() returns the child's exit code and
stderr; a sanitizer finding makes the child fail (halt_on_error).

    python -m oracle.sanitize asan|tsan
"""
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
OUT = HERE / "_san"

FLAGS = {
    "asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"],
    "tsan": ["-fsanitize=thread"],
}
RUNTIME = {"asan": ["libasan.so", "libubsan.so"], "tsan": ["libtsan.so"]}
OPTIONS = {
    "asan": {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
             "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"},
    "tsan": {"TSAN_OPTIONS": "halt_on_error=1:report_signal_unsafe=0"},
}

WORKLOAD = r"""
import numpy as np
import oracle
from gtpower_amd import abi, synthetic as S
d = S.make_design(600, dff_fraction=0.15, clock_ns=2.0, seed=5, threads=4)
tr = d.trace(300, toggle_p=0.2, seed=6, threads=4)
a = oracle.run(d.library, d, tr, threads=4)
b = oracle.run(d.library, d, tr, threads=1)
assert np.array_equal(a["gate_toggle_x2"], b["gate_toggle_x2"])
assert np.allclose(a["cycle_power"], b["cycle_power"], rtol=1e-12, atol=0)
text = S.vcd_text(tr, buses=[(10, 8)])
v = oracle.vcd_parse(text)
assert len(v["times"]) >= tr.n_event  # bus bits repeat values when another bit changes
print("sanitized workload ok", int(a["gate_toggle_x2"].sum()))
"""


# A planted mistake for each sanitizer, to show that the harness catches one:
# a heap read one past the end (asan), two threads adding to one int (tsan).
PLANTED = r"""
#include <thread>
#include <vector>
extern "C" int planted_overflow(int n) {
  std::vector<int> v(n, 1);
  int* p = v.data();
  volatile int i = n;
  return p[i];
}
extern "C" int planted_race(int n) {
  int x = 0;
  std::thread a([&] { for (int i = 0; i < n; ++i) x += 1; });
  std::thread b([&] { for (int i = 0; i < n; ++i) x += 1; });
  a.join();
  b.join();
  return x;
}
"""


def run_planted(kind, timeout=300):
    """Runs the planted mistake of `kind` under the same preloaded runtime;
    returns (exit code, stderr): a working harness fails with a report."""
    OUT.mkdir(exist_ok=True)
    src = OUT / "planted.cpp"
    src.write_text(PLANTED)
    so = OUT / f"libplanted_{kind}.so"
    subprocess.run(["g++", "-O0", "-g", "-std=c++17", "-fPIC", "-shared", "-pthread", *FLAGS[kind], str(src),
                    "-o", str(so)], check=True)
    fn = "planted_overflow" if kind == "asan" else "planted_race"
    code = f"import ctypes; L = ctypes.CDLL({str(so)!r}); print(L.{fn}(100000))"
    p = subprocess.run([sys.executable, "-c", code], env=_env(kind), capture_output=True, text=True,
                       timeout=timeout)
    return p.returncode, p.stderr


def _env(kind):
    pre = [subprocess.run(["gcc", f"-print-file-name={r}"], capture_output=True, text=True, check=True)
           .stdout.strip() for r in RUNTIME[kind]]
    env = dict(os.environ)
    env.update(OPTIONS[kind])
    env["LD_PRELOAD"] = " ".join(pre)
    env["PYTHONPATH"] = str(REPO)
    return env


def build(kind):
    OUT.mkdir(exist_ok=True)
    common = ["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-pthread", *FLAGS[kind]]
    outs = {"oracle": OUT / f"liboracle_{kind}.so", "synth": OUT / f"libgtpower_synth_{kind}.so"}
    for cmd in ([*common, str(HERE / "oracle.cpp"), str(HERE / "vcd.cpp"), "-o", str(outs["oracle"])],
                [*common, f"-I{REPO / 'include'}", str(REPO / "gtpower_amd" / "synth" / "synth.cpp"),
                 "-o", str(outs["synth"])]):
        print("+", " ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return outs


def run(kind, timeout=600):
    outs = build(kind)
    env = _env(kind)
    env["GTPOWER_ORACLE_SO"] = str(outs["oracle"])
    env["GTPOWER_SYNTH_SO"] = str(outs["synth"])
    p = subprocess.run([sys.executable, "-c", WORKLOAD], env=env, capture_output=True, text=True,
                       timeout=timeout, cwd=str(REPO))
    return p.returncode, p.stdout, p.stderr


if __name__ == "__main__":
    rc, out, err = run(sys.argv[1] if len(sys.argv) > 1 else "asan")
    print(out)
    print(err[-4000:], file=sys.stderr)
    sys.exit(rc)
