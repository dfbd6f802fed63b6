"""Hand-built libraries and one-gate designs for known-answer tests.

`const_library` replaces the internal_power groups of one cell of the bundled
sky130 library by groups with constant (one-point) tables, so that the energy
a rule selects names the group it came from.  Constants are in joules.
"""
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
