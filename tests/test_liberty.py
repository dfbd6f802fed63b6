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
