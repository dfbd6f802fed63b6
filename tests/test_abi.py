"""The C-ABI library loads and exports every function include/gtpower.h declares (CPU)."""
import re
import subprocess
from pathlib import Path

import pytest

from gtpower_amd import abi

REPO = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (REPO / "include" / "gtpower.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gtp_[a-z_0-9]+)\s*\(", text)))


def test_engine_exports_every_declared_symbol():
    so = abi.ENGINE_SO
    if not so.exists():
        pytest.fail(f"{so} missing: run __graft_entry__.build()")
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, f"not exported: {missing}"
    assert len(declared_functions()) >= 12


def test_engine_loads_without_device():
    lib = abi.load_engine()
    o = abi.GtpOptions()
    lib.gtp_options_init(o)
    assert o.device == -1


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof of the ABI structs, compiled with gcc, equal the ctypes mirror."""
    import ctypes as C
    src = tmp_path / "probe.c"
    structs = ["gtp_options", "gtp_table", "gtp_internal_power", "gtp_library", "gtp_design", "gtp_trace",
               "gtp_result", "gtp_schedule_info", "gtp_vcd_info", "gtp_vcd_var_info"]
    body = "\n".join(f'printf("%zu\\n", sizeof({s}));' for s in structs)
    src.write_text(f'#include <stdio.h>\n#include "gtpower.h"\nint main(void){{{body} return 0;}}\n')
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", str(REPO / "include"), str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    mirrors = [abi.GtpOptions, abi.GtpTable, abi.GtpInternalPower, abi.GtpLibrary, abi.GtpDesign, abi.GtpTrace,
               abi.GtpResult, abi.GtpScheduleInfo, abi.GtpVcdInfo, abi.GtpVcdVarInfo]
    assert sizes == [C.sizeof(m) for m in mirrors]


def test_reference_cycles_per_thread_matches_oracle():
    import numpy as np
    import oracle
    from gtpower_amd.power import reference_cycles_per_thread
    rng = np.random.default_rng(3)
    ev = rng.integers(0, 200000, size=300)
    for n_cycle in (1, 7, 1000, 100000):
        got = reference_cycles_per_thread(ev, n_cycle)
        want = [oracle.cycles_per_thread(int(e), n_cycle) for e in ev]
        assert got.tolist() == want


def test_writers_format(tmp_path):
    import numpy as np
    import ctypes as C
    lib = abi.load_engine()
    gp = np.arange(10, dtype=np.float64).reshape(5, 2) * 1e-6
    names = b"u1\0u2\0"
    cells = b"sky130_fd_sc_hd__inv_1\0sky130_fd_sc_hd__buf_1\0\0"
    gc = np.array([1, 0], np.int32)
    p = tmp_path / "cuda_cell_power.txt"
    assert lib.gtp_write_cell_power(str(p).encode(), 2, abi.ptr(np.ascontiguousarray(gp.reshape(-1)), C.c_double),
                                    names, abi.ptr(gc, C.c_int32), cells) == 0
    lines = p.read_text().splitlines()
    assert lines[0].split("\t") == ["cell", "Internal_Power", "Glitch_Internal_Power", "Switching_Power",
                                    "Glitch_Switching_Power", "Leakage_Power", "Total_Power", "Standard_Cell"]
    row = lines[1].split("\t")
    assert row[0] == "u1" and row[-1] == "sky130_fd_sc_hd__buf_1"
    assert float(row[6]) == pytest.approx(sum(float(v) for v in row[1:6]), rel=1e-6)
    cp = np.arange(15, dtype=np.float64) * 1e-3
    q = tmp_path / "cuda_per_cycle_waveform.txt"
    assert lib.gtp_write_per_cycle(str(q).encode(), b"top", 3, abi.ptr(cp, C.c_double)) == 0
    t = q.read_text().splitlines()
    assert t[0] == "Per cycle leakage power: " and t[-1] == "End of per cycle power"
    assert t[1].split("\t")[0] == "top" and len(t) == 11
