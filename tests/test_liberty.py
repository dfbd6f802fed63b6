"""The Liberty reader (gtpower_amd/liberty.py) and the flattened library it
feeds: the reference's sky130 tt data (an 8-cell excerpt of its
examples/sky130hd_tt.lib) reproduces the bundled tables bit for bit; a
hand-written library pins units, axis order by template variables, PVT
scaling (liberty/Liberty.cc:280-316) and leakage states; the reference's
ASAP7 test library (test/asap7_small.lib.gz) is read with its own units."""
import gzip
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
from gtpower_amd import abi
from gtpower_amd import liberty as L
from gtpower_amd.library import CellLibrary

GOLD = Path(__file__).resolve().parent / "golden"

HAND = """
library (hand) {
  time_unit : "1ps" ;
  capacitive_load_unit (1, ff) ;
  leakage_power_unit : "1pW";
  voltage_unit : "1V";
  nom_process : 1; nom_temperature : 25; nom_voltage : 0.8;
  k_process_internal_power : 0.5;
  k_temp_internal_power : 0.001;
  k_volt_internal_power : -0.5;
  operating_conditions (hot) { process : 1.2; temperature : 125; voltage : 0.9; }
  default_operating_conditions : hot;
  scaling_factors (cellk) { k_temp_internal_power : 0.002; }
  power_lut_template (cap_by_slew) {
    variable_1 : total_output_net_capacitance;
    variable_2 : input_transition_time;
    index_1 ("1, 2");
    index_2 ("10, 20, 40");
  }
  power_lut_template (slew_only) { variable_1 : input_transition_time; index_1 ("10, 20"); }
  cell (INVT) {
    area : 1
    cell_leakage_power : 3;
    leakage_power () { when : "A"; value : 5; }
    leakage_power () { when : "!A"; value : 7; }
    pin (A) { direction : input; capacitance : 1.5;
      internal_power () {
        rise_power (slew_only) { values ("1, 2"); }
        fall_power (slew_only) { values ("3, 4"); }
      }
    }
    pin (Y) { direction : output; function : "!A";
      internal_power () { related_pin : "A";
        rise_power (cap_by_slew) { values ("1, 2, 3", \\
                                           "4, 5, 6"); }
        fall_power (cap_by_slew) { index_1 ("1, 3"); values ("7, 8, 9", "10, 11, 12"); }
      }
      timing () { related_pin : "A"; timing_sense : negative_unate;
        cell_rise (scalar) { values ("5"); } cell_fall (scalar) { values ("6"); }
        rise_transition (scalar) { values ("7"); } fall_transition (scalar) { values ("8"); }
      }
    }
  }
  cell (INVK) {
    scaling_factors : cellk ;
    cell_leakage_power : 3;
    pin (A) { direction : input; capacitance : 1.5; }
    pin (Y) { direction : output; function : "!A";
      internal_power () { related_pin : "A";
        rise_power (cap_by_slew) { values ("1, 2, 3", "4, 5, 6"); }
        fall_power (cap_by_slew) { values ("1, 2, 3", "4, 5, 6"); }
      }
    }
  }
}
"""

S_T = 1.1 * 1.1 * 0.95        # (1 + 0.5*0.2)(1 + 0.001*100)(1 - 0.5*0.1)
S_K = 1.2                      # cell factors: (1 + 0.002*100)


def test_sky130_excerpt_matches_bundled_tables():
    doc = L.read_liberty(GOLD / "sky130hd_tt_8cells.lib")
    assert doc["units_si"] == {"time": 1e-9, "leakage_power": 1e-9, "voltage": 1.0, "capacitance": 1e-12,
                               "energy": 1e-12}
    a = CellLibrary(doc=doc, cells=list(doc["cells"]))
    b = CellLibrary()
    for k in ("table_data", "leak_values", "cell_leak_default", "table_array", "ipower_array", "cell_clock_mask",
              "cell_n_pin", "cell_n_in", "cell_class", "cell_leak_offset"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert all(c["internal_power_scale"] == 1.0 for c in doc["cells"].values())


def test_hand_library_known_answers():
    doc = L.read_liberty(HAND)
    u = doc["units_si"]
    assert u["time"] == 1e-12 and u["capacitance"] == 1e-15 and u["leakage_power"] == 1e-12
    assert u["energy"] == pytest.approx(1e-15, rel=1e-15)
    assert doc["cells"]["INVT"]["internal_power_scale"] == pytest.approx(S_T, rel=1e-12)
    assert doc["cells"]["INVK"]["internal_power_scale"] == pytest.approx(S_K, rel=1e-12)
    lib = CellLibrary(doc=doc, cells=["INVT", "INVK"])
    c = lib.cells[0]
    assert c["pins"] == ["A", "Y"] and c["pin_cap"][0] == pytest.approx(1.5e-15)
    # leakage by state (bit 0 = A): A=1 -> 5 pW, A=0 -> 7 pW, unscaled
    leak = lib.leak_values[lib.cell_leak_offset[0]:lib.cell_leak_offset[1]]
    assert np.allclose(leak, [7e-12, 5e-12, 7e-12, 5e-12], rtol=1e-15)
    assert lib.cell_leak_default[1] == pytest.approx(3e-12)
    # the output table: axis 1 is the load, axis 2 the slew, as its template says
    ip = [x for x in lib.ipower_array if x["cell"] == 0 and x["pin"] == 1][0]
    t = lib.table_array[ip["rise_table"]]
    assert (t["var1"], t["var2"]) == (abi.AXIS_CAP, abi.AXIS_SLEW)
    d = lib.table_data
    assert np.allclose(d[t["axis1"]:t["axis1"] + 2], [1e-15, 2e-15], rtol=1e-15)
    assert np.allclose(d[t["axis2"]:t["axis2"] + 3], [10e-12, 20e-12, 40e-12], rtol=1e-15)
    assert np.allclose(d[t["values"]:t["values"] + 6], np.arange(1, 7) * 1e-15 * S_T, rtol=1e-12)
    tf = lib.table_array[ip["fall_table"]]  # its own index_1 replaces the template's
    assert np.allclose(d[tf["axis1"]:tf["axis1"] + 2], [1e-15, 3e-15], rtol=1e-15)
    # the oracle's lookup names the axes by var: at slew 15 ps, load 1.5 fF the
    # bilinear value is 3 (rows 2.5, 3.5, 4.5 at 1.5 fF; 3.0 at 15 ps)

    assert oracle.table_value(lib, int(ip["rise_table"]), 15e-12, 1.5e-15) == pytest.approx(3e-15 * S_T, rel=1e-12)
    # the input table is 1-D in slew
    ipa = [x for x in lib.ipower_array if x["cell"] == 0 and x["pin"] == 0][0]
    assert oracle.table_value(lib, int(ipa["fall_table"]), 15e-12, 0.0) == pytest.approx(3.5e-15 * S_T, rel=1e-12)
    # nominal conditions: no scaling
    nom = L.read_liberty(HAND, operating_conditions={"process": 1, "temperature": 25, "voltage": 0.8})
    assert all(c["internal_power_scale"] == 1.0 for c in nom["cells"].values())


def _text_numbers(s):
    return [float(x) for x in re.findall(r"[-+0-9.eE]+", s)]


def test_asap7_library():
    doc = L.read_liberty(GOLD / "asap7_small.lib.gz")
    u = doc["units_si"]
    assert (u["time"], u["capacitance"], u["leakage_power"], u["voltage"]) == (1e-12, 1e-15, 1e-12, 1.0)
    assert doc["nom"]["voltage"] == 0.7 and doc["operating_conditions"]["voltage"] == 0.7
    names = list(doc["cells"])
    assert names == ["BUFx2_ASAP7_75t_R", "AND2x2_ASAP7_75t_R", "DFFHQx4_ASAP7_75t_R"]
    lib = CellLibrary(doc=doc, cells=names)
    assert list(lib.cell_n_pin) == [2, 3, 3] and list(lib.cell_class) == [abi.CLASS_COMB, abi.CLASS_COMB,
                                                                           abi.CLASS_SEQ]
    # spot check against the text: the first rise_power table of BUFx2's Y
    text = gzip.decompress((GOLD / "asap7_small.lib.gz").read_bytes()).decode()
    cell = text[text.index("cell (BUFx2_ASAP7_75t_R)"):text.index("cell (AND2x2_ASAP7_75t_R)")]
    blk = cell[cell.index("rise_power (power_template_7x7_x1)"):]
    v0 = blk.index("values")
    vals = _text_numbers(blk[v0 + len("values"):blk.index(");", v0)])
    ip = [x for x in lib.ipower_array if x["cell"] == 0 and x["pin"] == 1][0]
    t = lib.table_array[ip["rise_table"]]
    got = lib.table_data[t["values"]:t["values"] + 49]
    assert np.allclose(got, np.array(vals[:49]) * 1e-15, rtol=1e-15)
    assert (t["var1"], t["var2"]) == (abi.AXIS_SLEW, abi.AXIS_CAP)
    # leakage as the reference builds it (power/TimeBasedPower.cc:258-301):
    # groups in Liberty order overwrite the states their `when` matches, a
    # group without `when` matches all, related_pg_pin is not read
    # (liberty/LibertyReader.cc:5236-5240 keeps it for internal power only);
    # BUFx2's last group is an unconditional VSS 0, so every state is 0
    groups = re.findall(r"leakage_power \(\) \{(.*?)\}", cell[:cell.index("pin (Y)")], flags=re.S)
    assert len(groups) == 6 and "when" not in groups[-1]
    assert float(re.search(r"value : ([-+0-9.eE]+)", groups[-1]).group(1)) == 0.0
    leak = lib.leak_values[lib.cell_leak_offset[0]:lib.cell_leak_offset[1]]
    assert np.array_equal(leak, np.zeros(4))


# ---------------------------------------------------------------------------
# Cells of 7 and 8 pins (sky130 mux4_1, sdfbbp_1, sdfrbp_1, a222oi_1)
# ---------------------------------------------------------------------------
def _text_cells(path):
    """name -> (pin names inputs first then outputs, [(value, when)], cell_leakage_power),
    read from the Liberty text with regular expressions."""
    text = path.read_text()
    out = {}
    for m in re.finditer(r'^    cell \("sky130_fd_sc_hd__([^"]+)"\) \{\n(.*?)^    \}\n', text, re.M | re.S):
        body = m.group(2)
        pins = re.findall(r'^        pin \("([^"]+)"\) \{\n(.*?)^        \}\n', body, re.M | re.S)
        ins = [n for n, b in pins if 'direction : "input"' in b]
        outs = [n for n, b in pins if 'direction : "output"' in b]
        groups = [(float(v), w) for v, w in re.findall(
            r'leakage_power \(\) \{\s*value : ([-0-9.eE]+);\s*when : "([^"]*)";\s*\}', body)]
        default = float(re.search(r'^        cell_leakage_power : ([-0-9.eE]+);', body, re.M).group(1))
        out[m.group(1)] = (ins + outs, groups, default)
    return out


def test_large_cells_flatten_from_the_text():
    from hand_libs import LARGE_CELLS, sky130_library
    lib = sky130_library(large_only=True)
    text = _text_cells(GOLD / "sky130hd_tt_large_cells.lib")
    assert list(text) == list(LARGE_CELLS) == lib.cell_names
    want_pins = {"mux4_1": (7, 6), "sdfbbp_1": (8, 6), "sdfrbp_1": (7, 5), "a222oi_1": (7, 6)}
    for ci, name in enumerate(lib.cell_names):
        pins, groups, default = text[name]
        c = lib.cells[ci]
        assert c["pins"] == pins
        assert (c["n_pin"], c["n_in"]) == want_pins[name]
        seq = name.startswith("sdf")
        assert lib.cell_class[ci] == (abi.CLASS_SEQ if seq else abi.CLASS_COMB)
        assert lib.cell_clock_mask[ci] == (1 if seq else 0)          # CLK is input 0
        leak = lib.leak_values[lib.cell_leak_offset[ci]:lib.cell_leak_offset[ci + 1]]
        assert len(leak) == 1 << len(pins)
        named = set()
        for v, when in groups:
            # a conjunction of pin literals: every state that agrees with each
            # literal (pins it does not name take both values)
            lits = [(pins.index(x.lstrip("!")), x[0] != "!") for x in when.split("&")]
            states = [s for s in range(len(leak)) if all(bool((s >> k) & 1) == b for k, b in lits)]
            assert len(states) == 1 << (len(pins) - len(lits))
            assert np.all(leak[states] == v * 1e-9), (name, when)
            assert named.isdisjoint(states)
            named.update(states)
        rest = [s for s in range(len(leak)) if s not in named]
        assert np.all(leak[rest] == default * 1e-9)
        assert (name in ("mux4_1", "a222oi_1")) == (len(rest) == 0)
    # internal power of cells over 6 pins: unconditional groups only
    assert np.all(lib.ipower_array["when_truth"] == np.uint64(2**64 - 1))
    from gtpower_amd.library import when_truth
    with pytest.raises(ValueError):
        when_truth("S0", lib.cells[0]["pins"])


def _mux4_gate(lib, waves, t_max, slew=40e-12, cap=6e-15):
    from gtpower_amd import synthetic as S
    c = lib.index("mux4_1")
    d = S.ArrayDesign(lib, [c], list(range(7)), np.full(14, slew), [0.0] * 6 + [cap],
                      np.full(6 * 4, np.nan), 7)
    return d, S.trace_from_lists(waves, t_max, 1e-12, 1e-9)


@pytest.mark.parametrize("state", [0, 0b0010000, 0b1000101, 0b1111111])
def test_large_cell_oracle_leakage_known_answers(state):
    """A mux4_1 held in one state for the whole trace: its leakage is the
    Liberty value of that state (state index bit p = pin p, inputs first)."""
    from hand_libs import sky130_library
    lib = sky130_library(large_only=True)
    waves = [[(0, (state >> p) & 1)] for p in range(7)]
    d, tr = _mux4_gate(lib, waves, 4000)
    r = oracle.run(lib, d, tr, threads=1)
    pins, groups, _ = _text_cells(GOLD / "sky130hd_tt_large_cells.lib")["mux4_1"]
    want = [v for v, w in groups
            if all((x[0] != "!") == bool((state >> pins.index(x.lstrip("!"))) & 1) for x in w.split("&"))]
    assert len(want) == 1
    assert r["gate_power"][4, 0] == pytest.approx(want[0] * 1e-9, rel=1e-12)
    # per-cycle rows: leakage, internal, glitch internal, switching, glitch switching
    assert np.allclose(r["cycle_power"][0], want[0] * 1e-9, rtol=1e-12, atol=0)
    assert r["gate_power"][:4, 0].tolist() == [0, 0, 0, 0]


def test_large_cell_oracle_input_energy_in_a_high_state():
    """S0 of a mux4_1 rises at 100 while X is 1 (state index >= 64): the
    energy is S0's rise_power table (1-D in slew) at S0's rise slew; the
    leakage switches state at 100."""
    from hand_libs import sky130_library
    lib = sky130_library(large_only=True)
    waves = [[(0, 1)], [(0, 0)], [(0, 1)], [(0, 0)], [(0, 0), (100, 1)], [(0, 0)], [(0, 1)]]
    d, tr = _mux4_gate(lib, waves, 2000, slew=40e-12)
    r = oracle.run(lib, d, tr, threads=1)
    ip = [x for x in lib.ipower_array if x["cell"] == 0 and x["pin"] == 4]
    assert len(ip) == 1 and ip[0]["related_pin"] == -1
    # S0 rise_power values at index 40 ps, linear between 23.05 and 53.13 ps, in pJ
    ax = [0.0230506, 0.0531329]
    ys = [0.0033720, 0.0030965]
    e = (ys[0] + (ys[1] - ys[0]) * (0.040 - ax[0]) / (ax[1] - ax[0])) * 1e-12
    assert r["gate_power"][0, 0] == pytest.approx(e / 2e-9, rel=1e-9)
    assert r["cycle_power"][1].tolist() == pytest.approx([e / 1e-9, 0.0], rel=1e-9)
    assert r["gate_toggle_x2"][0] == 2
    leak = lib.leak_values[lib.cell_leak_offset[0]:lib.cell_leak_offset[1]]
    s0, s1 = 0b1000101, 0b1010101
    assert r["gate_power"][4, 0] == pytest.approx((leak[s0] * 100 + leak[s1] * 1900) / 2000, rel=1e-12)


def test_sixteen_pin_cell_leakage_and_energy_known_answers():
    """A hand-written 16-pin cell (15 inputs I0..I14, output Y): leakage by
    the state of all 16 pins, including states >= 2^15 (Y = 1), and the
    input energy of I3 from its 1-D table."""
    from gtpower_amd import synthetic as S
    from hand_libs import wide_cell_liberty
    doc = L.read_liberty(wide_cell_liberty(15, seed=1))
    lib = CellLibrary(doc=doc)
    assert lib.cell_n_pin[0] == 16 and lib.cell_n_in[0] == 15
    leak = lib.leak_values[lib.cell_leak_offset[0]:lib.cell_leak_offset[1]]
    assert len(leak) == 1 << 16
    # groups in Liberty order: I0&!I1 -> 2.5, I14&Y -> 3.5, !I0&!I1&!I2 -> 0.75, else 1.5 nW
    hi = (1 << 14) | (1 << 15)
    want = {0b1: 2.5, hi | 0b100: 3.5, hi | 0b1: 3.5,          # the later group overwrites
            hi: 0.75, 0: 0.75, 1 << 15: 0.75, 0b10: 1.5, 0b1010: 1.5}
    for s, v in want.items():
        assert leak[s] == pytest.approx(v * 1e-9, rel=1e-15), bin(s)
    # one gate: I3 rises at 500 with Y = 1, I14 = 1, I0 = 1; all else 0
    start = [0] * 16
    start[0] = start[14] = start[15] = 1
    waves = [[(0, v)] for v in start]
    waves[3].append((500, 1))
    d = S.ArrayDesign(lib, [0], list(range(16)), np.full(32, 0.1e-9), [0.0] * 15 + [3e-15],
                      np.full(15 * 4, np.nan), 16)
    tr = S.trace_from_lists(waves, 1000, 1e-12, 1e-9)
    r = oracle.run(lib, d, tr, threads=1)
    s0 = 1 | (1 << 14) | (1 << 15)
    assert r["gate_power"][4, 0] == pytest.approx((leak[s0] * 500 + leak[s0 | 8] * 500) / 1000, rel=1e-12)
    # the Liberty value at the grid point 0.1 ns, in pJ (1 V x 1 pF)
    e = doc["cells"]["WIDE16"]["pins"][3]["internal_power"][0]["rise_power"]["values"][1] * 1e-12
    assert r["gate_power"][0, 0] == pytest.approx(e / 1e-9, rel=1e-12)
