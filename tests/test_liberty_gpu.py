"""The engine on a library read by gtpower_amd.liberty with every 2-D power
table stored the other way round (template variable_1 = load, variable_2 =
slew) and scaled to operating conditions (factor 1.25): internal power is
1.25x that of the bundled library, everything else unchanged, and the oracle
agrees on the transposed library (GPU).  Tolerance: 1e-6 of each value's
magnitude scale (as in test_engine_gpu)."""
import copy
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from gtpower_amd import abi
from gtpower_amd import liberty as L
from gtpower_amd import synthetic as S
from gtpower_amd.library import CellLibrary
from gtpower_amd.power import HipPower

sys.path.insert(0, str(Path(__file__).resolve().parent))
from test_engine_gpu import compare  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
PREFIX = "sky130_fd_sc_hd__"


def transposed_scaled(scale):
    doc = L.read_liberty(GOLD / "sky130hd_tt_8cells.lib")
    doc = copy.deepcopy(doc)
    n2d = 0
    for c in doc["cells"].values():
        c["internal_power_scale"] = scale
        for p in c["pins"]:
            for ip in p["internal_power"]:
                for key in ("rise_power", "fall_power"):
                    t = ip[key]
                    if t is not None and "index_2" in t:
                        t["index_1"], t["index_2"] = t["index_2"], t["index_1"]
                        t["values"] = np.array(t["values"]).T.tolist()
                        t["vars"] = t["vars"][::-1]
                        n2d += 1
    assert n2d > 10
    doc["cells"] = {k[len(PREFIX):]: v for k, v in doc["cells"].items()}
    return doc


def test_transposed_scaled_library():
    base = CellLibrary()
    doc = transposed_scaled(1.25)
    lib = CellLibrary(doc=doc, cells=base.cell_names)
    assert any(t["var1"] == abi.AXIS_CAP for t in lib.table_array)
    d = S.make_design(1500, dff_fraction=0.15, clock_ns=2.0, seed=71)
    tr = d.trace(1200, toggle_p=0.2, seed=72)
    print("[liberty] bundled vs transposed+scaled library", flush=True)
    with HipPower() as e:
        e.set_design(base, d)
        e.read_power_activities(tr)
        a = e.power()
        e.set_design(lib, d)
        e.read_power_activities(tr)
        b = e.power()
    assert np.array_equal(a.gate_toggle_x2, b.gate_toggle_x2) and np.array_equal(a.gate_glitch, b.gate_glitch)
    tol = 1e-6
    for rows, ga, gb in ((slice(0, 2), a.gate_power, b.gate_power), (slice(1, 3), a.cycle_power, b.cycle_power)):
        want = 1.25 * ga[rows]
        scale = np.abs(want).max(axis=1, keepdims=True) + 1e-300
        assert np.all(np.abs(gb[rows] - want) <= tol * scale)
    assert np.allclose(b.gate_power[2:], a.gate_power[2:], rtol=1e-7, atol=0)
    assert np.allclose(b.cycle_power[[0, 3, 4]], a.cycle_power[[0, 3, 4]], rtol=1e-7, atol=0)
    print("[liberty] oracle on the transposed library", flush=True)
    ref = oracle.run(lib, d, tr)
    compare(b, ref, "transposed + scaled library")
