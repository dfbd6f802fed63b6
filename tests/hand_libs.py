"""Hand-built libraries and one-gate designs for known-answer tests.

`const_library` replaces the internal_power groups of one cell of the bundled
sky130 library by groups with constant (one-point) tables, so that the energy
a rule selects names the group it came from.  Constants are in joules.
"""
from pathlib import Path

import numpy as np

from gtpower_amd import abi
from gtpower_amd import synthetic as S
from gtpower_amd.library import CellLibrary, when_truth


def const_library(cell, groups):
    """A CellLibrary whose `cell` has exactly `groups`: a list, in Liberty
    order, of (pin, when, related_pin, rise_energy, fall_energy); pin and
    related_pin are pin names (related_pin None: no related pin)."""
    lib = CellLibrary()
    c = lib.index(cell)
    pins = lib.cells[c]["pins"]
    keep = [ip for ip in lib.ipowers if ip[0] != c]
    new = []
    for pin, when, rel, er, ef in groups:
        r = lib._add_table({"index_1": [0.0], "values": [er]}, abi.AXIS_SLEW, abi.AXIS_CAP, 1.0, 1.0, 1.0)
        f = lib._add_table({"index_1": [0.0], "values": [ef]}, abi.AXIS_SLEW, abi.AXIS_CAP, 1.0, 1.0, 1.0)
        new.append((c, pins.index(pin), pins.index(rel) if rel else -1, r, f, when_truth(when, pins)))
    lib.ipowers = keep + new
    lib._finish()
    return lib


def nand2_gate(lib, arcs, nets=(0, 1, 2), n_net=3, slew=50e-12, cap=4e-15):
    """One nand2_1 (pins A, B, Y) with the given arc delays
    {(input, from_rf, to_rf): seconds}; every other arc is absent (NaN)."""
    c = lib.index("nand2_1")
    arc = np.full(8, np.nan)
    for (i, fr, to), d in arcs.items():
        arc[(i * 2 + fr) * 2 + to] = d
    return S.ArrayDesign(lib, [c], list(nets), np.full(6, slew), [0.0, 0.0, cap], arc, n_net)


fJ = 1e-15
OUT_GROUPS = [
    ("Y", "!B", "A", 11 * fJ, 12 * fJ),
    ("Y", "B", "A", 13 * fJ, 14 * fJ),
    ("Y", "", "B", 15 * fJ, 16 * fJ),
]

# (name, groups, arcs, nets, waves): one nand2_1 on a 4000-tick trace of
# 1000-tick cycles (1 ps ticks)
WHEN_CASES = [
    ("input_pre_state",
     [("A", "B", None, 1 * fJ, 2 * fJ), ("A", "!B", None, 3 * fJ, 4 * fJ), ("B", "", None, 5 * fJ, 6 * fJ)],
     {}, (0, 1, -1),
     [[(0, 0), (100, 1), (2100, 0), (3100, 1)], [(0, 0), (100, 1), (3100, 0)], []]),
    ("input_first_match",
     [("A", "B", None, 1 * fJ, 2 * fJ), ("A", "", None, 7 * fJ, 8 * fJ)],
     {}, (0, 1, -1),
     [[(0, 0), (100, 1), (1100, 0), (3100, 1)], [(0, 1), (2050, 0)], []]),
    ("input_x_is_zero",
     [("A", "!B", None, 3 * fJ, 4 * fJ), ("A", "B", None, 1 * fJ, 2 * fJ)],
     {}, (0, 1, -1),
     [[(0, 0), (100, 1), (3100, 0)], [(0, 2), (2500, 1)], []]),
    ("related_tie", OUT_GROUPS, {(0, 0, 1): 40e-12, (1, 0, 1): 20e-12}, (0, 1, 2),
     [[(0, 0), (100, 1)], [(0, 0), (120, 1)], [(0, 1), (140, 0)]]),
    ("related_no_arc", OUT_GROUPS, {(1, 0, 1): 20e-12, (0, 1, 0): 30e-12}, (0, 1, 2),
     [[(0, 0), (100, 1)], [(0, 0), (50, 1)], [(0, 1), (140, 0)]]),
    ("related_earlier_cycle", OUT_GROUPS, {(0, 0, 1): 150e-12, (1, 0, 1): 100e-12}, (0, 1, 2),
     [[(0, 0), (990, 1)], [(0, 0), (1100, 1)], [(0, 1), (1140, 0)]]),
    ("output_first_match",
     [("Y", "B", "A", 13 * fJ, 14 * fJ), ("Y", "", "A", 19 * fJ, 20 * fJ)],
     {(0, 0, 1): 40e-12}, (0, 1, 2),
     [[(0, 0), (100, 1)], [(0, 1)], [(0, 1), (140, 0)]]),
    ("output_pre_state", OUT_GROUPS, {(0, 0, 1): 40e-12, (1, 0, 1): 20e-12}, (0, 1, 2),
     [[(0, 0), (100, 1)], [(0, 1), (140, 0)], [(0, 1), (140, 0)]]),
    ("output_default", OUT_GROUPS, {(0, 0, 1): 40e-12}, (0, 1, 2),
     [[(0, 0), (100, 1)], [(0, 1)], [(0, 1), (3500, 0)]]),
]


def when_case(name):
    """(library, design, trace) of WHEN_CASES entry `name`."""
    _, groups, arcs, nets, waves = next(c for c in WHEN_CASES if c[0] == name)
    lib = const_library("nand2_1", groups)
    d = nand2_gate(lib, arcs, nets=nets)
    tr = S.trace_from_lists(waves, 4000, 1e-12, 1e-9)
    return lib, d, tr


# ---------------------------------------------------------------------------
# Cells of 7 and 8 pins: the reference's sky130 tt data (tests/golden excerpts
# of examples/sky130hd_tt.lib) read with gtpower_amd.liberty
# ---------------------------------------------------------------------------
GOLD = Path(__file__).resolve().parent / "golden"
PREFIX = "sky130_fd_sc_hd__"
LARGE_CELLS = ("mux4_1", "sdfbbp_1", "sdfrbp_1", "a222oi_1")


def sky130_library(large_only=False):
    """The 8-cell and the large-cell excerpts in one CellLibrary (cell names
    without the sky130_fd_sc_hd__ prefix; the large cells last)."""
    from gtpower_amd import liberty as L
    big = L.read_liberty(GOLD / "sky130hd_tt_large_cells.lib")
    cells = {} if large_only else dict(L.read_liberty(GOLD / "sky130hd_tt_8cells.lib")["cells"])
    cells.update(big["cells"])
    doc = dict(big)
    doc["cells"] = {k[len(PREFIX):]: v for k, v in cells.items()}
    return CellLibrary(doc=doc)


def random_hand_design(lib, n_gate, n_primary, seed, large_share=0.5, p_unconnected=0.03, p_no_arc=0.1,
                       large_cells=LARGE_CELLS):
    """Gates drawn from `lib` (half of them, by default, from `large_cells`),
    each input on a random earlier net (clock pins on net 0, a few inputs
    unconnected), each output on a net of its own; random slews, loads and
    arc delays (some arcs absent)."""
    rng = np.random.default_rng(seed)
    large = [lib.index(c) for c in large_cells if c in lib.cell_names]
    small = [k for k in range(len(lib.cells)) if k not in large]
    gate_cell, pin_net, arcs = [], [], []
    n_net = n_primary
    for _ in range(n_gate):
        pool = large if (not small or rng.random() < large_share) else small
        c = int(rng.choice(pool))
        cell = lib.cells[c]
        n, nin = cell["n_pin"], cell["n_in"]
        for p in range(nin):
            if (cell["clock_mask"] >> p) & 1:
                pin_net.append(0)
            elif rng.random() < p_unconnected:
                pin_net.append(-1)
            else:
                pin_net.append(int(rng.integers(1, n_net)))
        for _ in range(n - nin):
            pin_net.append(n_net)
            n_net += 1
        a = rng.uniform(10e-12, 300e-12, size=(n - nin) * nin * 4)
        a[rng.random(a.size) < p_no_arc] = np.nan
        arcs.append(a)
        gate_cell.append(c)
    P = len(pin_net)
    return S.ArrayDesign(lib, gate_cell, pin_net, rng.uniform(5e-12, 120e-12, 2 * P),
                         rng.uniform(0.2e-15, 20e-15, P), np.concatenate(arcs), n_net,
                         pin_default_state=rng.choice([0, 1, 2], size=P, p=[0.3, 0.3, 0.4]))


def random_hand_trace(n_net, n_cycle, seed, T=1000, rate=0.4, p_x=0.05, p_silent=0.05):
    """Net 0 a clock (rise at 0, fall at T/2); every other net a random
    waveform: an initial value at time 0 (most nets), then about `rate`
    events per cycle at random ticks (some X values, some nets silent)."""
    rng = np.random.default_rng(seed)
    waves = [[e for c in range(n_cycle) for e in ((c * T, 1), (c * T + T // 2, 0))]]
    for _ in range(1, n_net):
        if rng.random() < p_silent:
            waves.append([])
            continue
        k = rng.poisson(rate * n_cycle)
        times = np.unique(rng.integers(1, n_cycle * T, size=k))
        if rng.random() < 0.8:
            times = np.concatenate([[0], times])
        vals = rng.integers(0, 2, size=len(times))
        vals[rng.random(len(times)) < p_x] = 2
        waves.append(list(zip(times.tolist(), vals.tolist())))
    return S.trace_from_lists(waves, n_cycle * T, 1e-12, T * 1e-12)


def wide_cell_liberty(n_in, seed=0, name=None, leak_scale=1.0):
    """Liberty text of one combinational cell of n_in inputs I0.. and output Y
    (n_in + 1 pins): random 1-D input and 2-D output power tables (one output
    group per related input), three leakage groups over a cell default (nW,
    times leak_scale)."""
    rng = np.random.default_rng(seed)
    name = name or f"WIDE{n_in + 1}"

    def vals(n):
        return ", ".join(f"{v:.6f}" for v in rng.uniform(0.0005, 0.01, n))

    pins = []
    for k in range(n_in):
        pins.append(f"""    pin (I{k}) {{ direction : input; capacitance : {rng.uniform(0.001, 0.004):.6f};
      internal_power () {{ rise_power (s1) {{ values ("{vals(3)}"); }} fall_power (s1) {{ values ("{vals(3)}"); }} }}
    }}""")
    groups = []
    for k in range(n_in):
        groups.append(f"""      internal_power () {{ related_pin : "I{k}";
        rise_power (s2) {{ values ("{vals(3)}", "{vals(3)}", "{vals(3)}"); }}
        fall_power (s2) {{ values ("{vals(3)}", "{vals(3)}", "{vals(3)}"); }}
      }}""")
    func = " | ".join(f"(I{k}&I{k + 1})" for k in range(0, n_in - 1, 2)) or "!I0"
    leak = [("I0&!I1", 2.5), (f"I{n_in - 1}&Y", 3.5), ("!I0&!I1&!I2", 0.75)]
    leak = [g for g in leak if n_in >= 3 or g[1] == 3.5]       # groups over existing pins only
    leak_txt = "\n".join(f'    leakage_power () {{ when : "{w}"; value : {v * leak_scale}; }}' for w, v in leak)
    return f"""library (wide) {{
  time_unit : "1ns"; capacitive_load_unit (1, pf); leakage_power_unit : "1nW"; voltage_unit : "1V";
  power_lut_template (s1) {{ variable_1 : input_transition_time; index_1 ("0.01, 0.1, 1"); }}
  power_lut_template (s2) {{ variable_1 : input_transition_time; variable_2 : total_output_net_capacitance;
    index_1 ("0.01, 0.1, 1"); index_2 ("0.001, 0.01, 0.1"); }}
  cell ({name}) {{
    cell_leakage_power : {1.5 * leak_scale};
{leak_txt}
{chr(10).join(pins)}
    pin (Y) {{ direction : output; function : "{func}";
{chr(10).join(groups)}
    }}
  }}
}}
"""


def wide_library(sizes=(11, 15)):
    """sky130 excerpts plus one hand-written cell per entry of `sizes` (its
    inputs; the cells have sizes + 1 pins, named WIDE12, WIDE16, ..., a
    repeated size WIDE<pins>_<entry>)."""
    from gtpower_amd import liberty as L
    lib = sky130_library()
    doc = lib.doc
    for k, n in enumerate(sizes):
        name = f"WIDE{n + 1}" if f"WIDE{n + 1}" not in doc["cells"] else f"WIDE{n + 1}_{k}"
        w = L.read_liberty(wide_cell_liberty(n, seed=k, name=name))
        doc["cells"].update(w["cells"])
    return CellLibrary(doc=doc)
