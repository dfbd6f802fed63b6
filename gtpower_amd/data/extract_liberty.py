"""Extract the power/timing data of a few sky130 cells from a Liberty file into JSON.

The engine never parses Liberty at run time: a design hands it flattened tables.
This script turns the eight sky130_fd_sc_hd cells that the synthetic designs use
(inv_1, buf_1, nand2_1, nor2_1, xor2_1, a21oi_1, mux2_1, dfxtp_1) into
``sky130hd_tt_subset.json`` next to this file.  The JSON holds data only (table
axes and values, pin directions and capacitances, `when` strings, functions);
it is the sky130 tt corner as published in the SkyWater PDK.

Usage:  python extract_liberty.py /path/to/sky130hd_tt.lib [out.json]

Liberty groups read here, by their meaning in the reference flow:
  - leakage_power{value,when} and cell_leakage_power: state-dependent leakage
    (reference: power/TimeBasedPower.cc:258-301 getLeakagePower).
  - pin/internal_power{related_pin,when,rise_power,fall_power}: internal energy
    tables (reference: power/TimeBasedPower.cc:341-555).
  - pin/timing{related_pin,timing_type,timing_sense,cell_rise,...}: NLDM delay
    and slew tables, used only by the synthetic design generator.
"""
import json
import re
import sys
from pathlib import Path

CELLS = ["inv_1", "buf_1", "nand2_1", "nor2_1", "xor2_1", "a21oi_1", "mux2_1", "dfxtp_1"]
PREFIX = "sky130_fd_sc_hd__"

_TOKEN = re.compile(r'"(?:[^"\\]|\\.)*"|[A-Za-z_0-9.\-+!&|\^*\[\]]+|[{}();:,]')


def _tokens(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = text.replace("\\\n", " ")
    return _TOKEN.findall(text)


class Group:
    def __init__(self, kind, args):
        self.kind = kind
        self.args = args
        self.attrs = {}
        self.complex = {}
        self.groups = []

    def sub(self, kind):
        return [g for g in self.groups if g.kind == kind]


def _parse_group_body(toks, i, group):
    # toks[i] is the token after '{'
    while i < len(toks):
        t = toks[i]
        if t == "}":
            return i + 1
        name = t
        i += 1
        if toks[i] == ":":
            # simple attribute: name : value ;
            j = i + 1
            vals = []
            while toks[j] != ";":
                vals.append(toks[j])
                j += 1
            group.attrs[name] = " ".join(vals).strip('"')
            i = j + 1
        elif toks[i] == "(":
            # complex attribute or group header: name ( args ) [{ ... } | ;]
            j = i + 1
            args = []
            cur = []
            depth = 0
            while not (toks[j] == ")" and depth == 0):
                if toks[j] == "(":
                    depth += 1
                elif toks[j] == ")":
                    depth -= 1
                if toks[j] == "," and depth == 0:
                    args.append(" ".join(cur))
                    cur = []
                else:
                    cur.append(toks[j])
                j += 1
            if cur:
                args.append(" ".join(cur))
            args = [a.strip('"') for a in args]
            j += 1
            if j < len(toks) and toks[j] == "{":
                child = Group(name, args)
                j = _parse_group_body(toks, j + 1, child)
                group.groups.append(child)
                i = j
            else:
                group.complex.setdefault(name, []).append(args)
                if j < len(toks) and toks[j] == ";":
                    j += 1
                i = j
        else:
            raise ValueError(f"unexpected token {toks[i]!r} after {name!r}")
    return i


def _floats(s):
    return [float(x) for x in s.replace('"', " ").replace(",", " ").split()]


def _table(g):
    if g is None:
        return None
    out = {"template": g.args[0] if g.args else ""}
    for key in ("index_1", "index_2"):
        if key in g.complex:
            out[key] = _floats(g.complex[key][0][0])
    vals = g.complex["values"][0]
    rows = [_floats(v) for v in vals]
    out["values"] = rows if len(rows) > 1 or "index_2" in out else rows[0]
    return out


def _one(g, kind):
    s = g.sub(kind)
    return s[0] if s else None


def extract_cell(cell):
    out = {"name": cell.args[0], "pins": [], "leakage_power": [],
           "cell_leakage_power": float(cell.attrs.get("cell_leakage_power", "nan"))}
    for lp in cell.sub("leakage_power"):
        out["leakage_power"].append({"value": float(lp.attrs["value"]),
                                     "when": lp.attrs.get("when", "")})
    ff = _one(cell, "ff")
    if ff is not None:
        out["ff"] = {"clocked_on": ff.attrs.get("clocked_on", ""),
                     "next_state": ff.attrs.get("next_state", ""),
                     "state": ff.args}
    for pin in cell.sub("pin"):
        p = {"name": pin.args[0], "direction": pin.attrs.get("direction", ""),
             "capacitance": float(pin.attrs.get("capacitance", "0")),
             "rise_capacitance": float(pin.attrs.get("rise_capacitance", pin.attrs.get("capacitance", "0"))),
             "fall_capacitance": float(pin.attrs.get("fall_capacitance", pin.attrs.get("capacitance", "0"))),
             "clock": pin.attrs.get("clock", "false") == "true",
             "function": pin.attrs.get("function", ""),
             "internal_power": [], "timing": []}
        for ip in pin.sub("internal_power"):
            p["internal_power"].append({
                "related_pin": ip.attrs.get("related_pin", ""),
                "when": ip.attrs.get("when", ""),
                "rise_power": _table(_one(ip, "rise_power")),
                "fall_power": _table(_one(ip, "fall_power")),
            })
        for tm in pin.sub("timing"):
            ttype = tm.attrs.get("timing_type", "combinational")
            if ttype not in ("combinational", "rising_edge", "falling_edge"):
                continue  # constraints (setup/hold/min_pulse_width) are not delay arcs
            p["timing"].append({
                "related_pin": tm.attrs.get("related_pin", ""),
                "timing_type": ttype,
                "timing_sense": tm.attrs.get("timing_sense", "non_unate"),
                "cell_rise": _table(_one(tm, "cell_rise")),
                "cell_fall": _table(_one(tm, "cell_fall")),
                "rise_transition": _table(_one(tm, "rise_transition")),
                "fall_transition": _table(_one(tm, "fall_transition")),
            })
        out["pins"].append(p)
    return out


def _cell_text(lines, start):
    depth = 0
    buf = []
    for k in range(start, len(lines)):
        line = lines[k]
        buf.append(line)
        depth += line.count("{") - line.count("}")
        if depth == 0 and k > start:
            break
    return "".join(buf)


def main(argv):
    lib_path = Path(argv[1])
    out_path = Path(argv[2]) if len(argv) > 2 else Path(__file__).with_name("sky130hd_tt_subset.json")
    lines = lib_path.read_text().splitlines(keepends=True)
    header = {}
    for line in lines[:400]:
        m = re.match(r'\s*(time_unit|voltage_unit|leakage_power_unit|capacitive_load_unit|nom_voltage)\s*[:(]\s*(.*?)\s*\)?;', line)
        if m:
            header[m.group(1)] = m.group(2).replace('"', "")
    vmap = {}
    for line in lines[:400]:
        m = re.match(r'\s*voltage_map\("(\w+)",\s*([0-9.eE+-]+)\)', line)
        if m:
            vmap[m.group(1)] = float(m.group(2))
    cells = {}
    for short in CELLS:
        name = PREFIX + short
        pat = f'cell ("{name}")'
        start = next(k for k, l in enumerate(lines) if pat in l)
        toks = _tokens(_cell_text(lines, start))
        # toks: cell ( name ) { ... }
        root = Group("root", [])
        _parse_group_body(toks + ["}"], 0, root)
        cells[short] = extract_cell(root.groups[0])
        cells[short]["source_line"] = start + 1
    doc = {"library": lib_path.name, "units": header, "voltage_map": vmap,
           "energy_unit": "pJ (voltage_unit^2 * capacitive_load_unit)",
           "cells": cells}
    out_path.write_text(json.dumps(doc, indent=1))
    print(f"wrote {out_path} ({out_path.stat().st_size} bytes)")


if __name__ == "__main__":
    main(sys.argv)
