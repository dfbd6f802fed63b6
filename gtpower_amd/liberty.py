"""Liberty reader for the power path: library text -> the cell description that
`CellLibrary` flattens for the engine.

The reference reads Liberty through OpenSTA (liberty/LibertyReader.cc) and
evaluates internal power as InternalPowerModel::power (liberty/InternalPower.cc:
145-229): the table's axes are matched to (input slew, load capacitance) by the
template's `variable_1/2/3` (axisValue), the value is scaled to the operating
conditions (TableModel::findValue with a Pvt, liberty/TableModel.cc:659-681;
LibertyLibrary::scaleFactor, liberty/Liberty.cc:280-316:
(1 + k_process * dP) (1 + k_temp * dT) (1 + k_volt * dV), cell scaling_factors
before library ones, dX from the library's nom_process / nom_temperature /
nom_voltage).  Leakage is read unscaled (power/TimeBasedPower.cc:258-301).

Here the text is parsed once on the host; `read_liberty` returns, per cell,
the pins with their internal_power groups (related pin, `when`, rise/fall
tables with each axis named SLEW or CAP and axes filled from the template),
the leakage groups and cell_leakage_power, the ff group and timing arcs (for
the synthetic generator), plus the library units (SI factors) and the
internal-power scale factor of each cell at the chosen operating conditions.
Tables of other kinds (CCS current, receiver capacitance, constraints) are
skipped.
"""
import gzip
import re
from pathlib import Path

from . import abi

_TOKEN = re.compile(r'"(?:[^"\\]|\\.)*"|[^\s{}();:,"]+|[{}();:,]|\n')


def _strip(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return text.replace("\\\r\n", " ").replace("\\\n", " ")


class Group:
    """A Liberty group: name(args) { simple : value; complex(args); subgroups }."""

    __slots__ = ("kind", "args", "attrs", "complex", "groups")

    def __init__(self, kind, args):
        self.kind = kind
        self.args = args
        self.attrs = {}
        self.complex = {}
        self.groups = []

    def sub(self, kind):
        return [g for g in self.groups if g.kind == kind]

    def one(self, kind):
        for g in self.groups:
            if g.kind == kind:
                return g
        return None


def _unquote(s):
    return s[1:-1] if len(s) >= 2 and s[0] == '"' and s[-1] == '"' else s


def parse_groups(text):
    """All top-level groups of a Liberty text (normally one `library`)."""
    # newlines are kept: a simple attribute ends at ';' or at the end of its line
    toks = _TOKEN.findall(_strip(text))
    root = Group("root", [])
    stack = [root]
    i, n = 0, len(toks)
    while i < n:
        t = toks[i]
        if t == "}":
            stack.pop()
            if not stack:
                raise ValueError("unbalanced '}'")
            i += 1
            continue
        if t in (";", "\n"):
            i += 1
            continue
        name = t
        i += 1
        while i < n and toks[i] == "\n":
            i += 1
        if i >= n:
            break
        if toks[i] == ":":
            j = i + 1
            vals = []
            while j < n and toks[j] not in (";", "}", "\n"):
                vals.append(toks[j])
                j += 1
            stack[-1].attrs[name] = _unquote(" ".join(vals))
            i = j + 1 if j < n and toks[j] in (";", "\n") else j
        elif toks[i] == "(":
            j = i + 1
            args, cur, depth = [], [], 0
            while j < n and not (toks[j] == ")" and depth == 0):
                if toks[j] == "\n":
                    j += 1
                    continue
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
            args = [_unquote(a) for a in args]
            j += 1
            while j < n and toks[j] == "\n":
                j += 1
            if j < n and toks[j] == "{":
                g = Group(name, args)
                stack[-1].groups.append(g)
                stack.append(g)
                i = j + 1
            else:
                stack[-1].complex.setdefault(name, []).append(args)
                i = j + 1 if j < n and toks[j] == ";" else j
        else:
            raise ValueError(f"unexpected token {toks[i]!r} after {name!r}")
    if len(stack) != 1:
        raise ValueError("unbalanced '{'")
    return root.groups


def _floats(s):
    return [float(x) for x in s.replace('"', " ").replace(",", " ").replace("\\", " ").split()]


_UNIT = {"s": 1.0, "ms": 1e-3, "us": 1e-6, "ns": 1e-9, "ps": 1e-12, "fs": 1e-15}
_POWER = {"w": 1.0, "mw": 1e-3, "uw": 1e-6, "nw": 1e-9, "pw": 1e-12, "fw": 1e-15}
_CAP = {"f": 1.0, "mf": 1e-3, "uf": 1e-6, "nf": 1e-9, "pf": 1e-12, "ff": 1e-15}
_VOLT = {"v": 1.0, "mv": 1e-3, "kv": 1e3}


def _unit(s, table, what):
    m = re.fullmatch(r"\s*([0-9.eE+-]*)\s*([A-Za-z]+)\s*", s)
    if not m or m.group(2).lower() not in table:
        raise ValueError(f"unsupported {what} unit {s!r}")
    k = float(m.group(1)) if m.group(1) else 1.0
    return k * table[m.group(2).lower()]


# template variables -> the engine's axis meaning (InternalPower.cc:215-229)
_AXIS_VAR = {"input_transition_time": abi.AXIS_SLEW, "input_net_transition": abi.AXIS_SLEW,
             "total_output_net_capacitance": abi.AXIS_CAP}

def _k_factors(attrs, prefix_type):
    return {p: float(attrs.get(f"k_{p}_{prefix_type}", 0.0)) for p in ("process", "temp", "volt")}


def _table(g, templates):
    """A power/timing table with axes from the group or its template and each
    axis named (SLEW / CAP), or None for a table over other variables."""
    if g is None:
        return None
    name = g.args[0] if g.args else ""
    tpl = templates.get(name, {})
    out = {"template": name}
    variables = tpl.get("variables", [])
    for key in ("index_1", "index_2", "index_3"):
        if key in g.complex:
            out[key] = _floats(g.complex[key][0][0])
        elif key in tpl:
            out[key] = list(tpl[key])
    if "index_3" in out:
        raise ValueError(f"3-D table {name!r} is not supported")
    rows = [_floats(v) for v in g.complex["values"][0]]
    two = "index_2" in out
    vals = [x for r in rows for x in r]
    if name in ("scalar", "") and not out.get("index_1"):
        out["index_1"] = [0.0]
        out["values"] = vals[:1]
        out["vars"] = [abi.AXIS_SLEW]
        return out
    n1 = len(out["index_1"])
    if two:
        n2 = len(out["index_2"])
        if len(vals) != n1 * n2:
            raise ValueError(f"table {name!r}: {len(vals)} values for {n1}x{n2} axes")
        out["values"] = [vals[i * n2:(i + 1) * n2] for i in range(n1)]
    else:
        if len(vals) != n1:
            raise ValueError(f"table {name!r}: {len(vals)} values for {n1} points")
        out["values"] = vals
    vs = variables[:2 if two else 1] or ["input_transition_time", "total_output_net_capacitance"][:2 if two else 1]
    if any(v not in _AXIS_VAR for v in vs):
        return None
    out["vars"] = [_AXIS_VAR[v] for v in vs]
    return out


def _pins(cell):
    """Pins of a cell, those inside bus / bundle groups included."""
    out = []
    for g in cell.groups:
        if g.kind == "pin":
            out += [(a, g) for a in g.args]
        elif g.kind in ("bus", "bundle"):
            out += [(a, p) for p in g.sub("pin") for a in p.args]
    return out


def read_liberty(source, cells=None, operating_conditions=None):
    """Read a Liberty library (path to .lib / .lib.gz, or the text).

    cells: names to keep (default: all); operating_conditions: the name of an
    operating_conditions group, or a dict(process, temperature, voltage), or
    None for the library's default_operating_conditions (no scaling when it
    has none)."""
    if isinstance(source, (bytes, bytearray)):
        text = bytes(source).decode("latin-1")
    elif isinstance(source, str) and "\n" in source:
        text = source
    else:
        p = Path(source)
        raw = p.read_bytes()
        text = (gzip.decompress(raw) if raw[:2] == b"\x1f\x8b" else raw).decode("latin-1")
    libs = [g for g in parse_groups(text) if g.kind == "library"]
    if not libs:
        raise ValueError("no library group")
    lib = libs[0]
    a = lib.attrs
    units = {
        "time": _unit(a.get("time_unit", "1ns"), _UNIT, "time"),
        "leakage_power": _unit(a.get("leakage_power_unit", "1nW"), _POWER, "leakage power"),
        "voltage": _unit(a.get("voltage_unit", "1V"), _VOLT, "voltage"),
    }
    cl = lib.complex.get("capacitive_load_unit")
    units["capacitance"] = float(cl[0][0]) * _CAP[cl[0][1].lower()] if cl else 1e-12
    # Liberty internal power tables are energies in voltage_unit^2 * capacitive_load_unit
    units["energy"] = units["voltage"] ** 2 * units["capacitance"]
    templates = {}
    for kind in ("lu_table_template", "power_lut_template"):
        for g in lib.sub(kind):
            t = {"variables": [g.attrs[f"variable_{k}"] for k in (1, 2, 3) if f"variable_{k}" in g.attrs]}
            for key in ("index_1", "index_2", "index_3"):
                if key in g.complex:
                    t[key] = _floats(g.complex[key][0][0])
            templates[g.args[0] if g.args else ""] = t
    nom = {"process": float(a.get("nom_process", 1.0)), "temperature": float(a.get("nom_temperature", 25.0)),
           "voltage": float(a.get("nom_voltage", 1.0))}
    opconds = {g.args[0]: {"process": float(g.attrs.get("process", nom["process"])),
                           "temperature": float(g.attrs.get("temperature", nom["temperature"])),
                           "voltage": float(g.attrs.get("voltage", nom["voltage"]))}
               for g in lib.sub("operating_conditions") if g.args}
    if operating_conditions is None:
        pvt = opconds.get(a.get("default_operating_conditions", ""))
    elif isinstance(operating_conditions, str):
        pvt = opconds[operating_conditions]
    else:
        pvt = dict(operating_conditions)
    lib_k = _k_factors(a, "internal_power")
    k_groups = {g.args[0]: _k_factors(g.attrs, "internal_power") for g in lib.sub("scaling_factors") if g.args}
    vmap = {args[0]: float(args[1]) for args in lib.complex.get("voltage_map", []) if len(args) >= 2}

    out_cells = {}
    want = set(cells) if cells is not None else None
    for cg in lib.sub("cell"):
        name = cg.args[0] if cg.args else ""
        if want is not None and name not in want:
            continue
        k = k_groups.get(cg.attrs.get("scaling_factors", ""), lib_k)
        scale = 1.0
        if pvt is not None:
            scale = ((1.0 + (pvt["process"] - nom["process"]) * k["process"])
                     * (1.0 + (pvt["temperature"] - nom["temperature"]) * k["temp"])
                     * (1.0 + (pvt["voltage"] - nom["voltage"]) * k["volt"]))
        c = {"name": name, "pins": [], "leakage_power": [],
             "cell_leakage_power": float(cg.attrs.get("cell_leakage_power", "nan")),
             "internal_power_scale": scale}
        for lp in cg.sub("leakage_power"):
            c["leakage_power"].append({"value": float(lp.attrs["value"]), "when": lp.attrs.get("when", "")})
        ff = cg.one("ff")
        if ff is not None:
            c["ff"] = {"clocked_on": ff.attrs.get("clocked_on", ""), "next_state": ff.attrs.get("next_state", ""),
                       "state": ff.args}
        for pname, pg in _pins(cg):
            cap = float(pg.attrs.get("capacitance", "0"))
            p = {"name": pname, "direction": pg.attrs.get("direction", ""), "capacitance": cap,
                 "rise_capacitance": float(pg.attrs.get("rise_capacitance", cap)),
                 "fall_capacitance": float(pg.attrs.get("fall_capacitance", cap)),
                 "clock": pg.attrs.get("clock", "false") == "true",
                 "function": pg.attrs.get("function", ""), "internal_power": [], "timing": []}
            for ip in pg.sub("internal_power"):
                p["internal_power"].append({
                    "related_pin": ip.attrs.get("related_pin", ""), "when": ip.attrs.get("when", ""),
                    "rise_power": _table(ip.one("rise_power"), templates),
                    "fall_power": _table(ip.one("fall_power"), templates)})
            for tm in pg.sub("timing"):
                ttype = tm.attrs.get("timing_type", "combinational")
                if ttype not in ("combinational", "rising_edge", "falling_edge"):
                    continue
                p["timing"].append({
                    "related_pin": tm.attrs.get("related_pin", ""), "timing_type": ttype,
                    "timing_sense": tm.attrs.get("timing_sense", "non_unate"),
                    "cell_rise": _table(tm.one("cell_rise"), templates),
                    "cell_fall": _table(tm.one("cell_fall"), templates),
                    "rise_transition": _table(tm.one("rise_transition"), templates),
                    "fall_transition": _table(tm.one("fall_transition"), templates)})
            c["pins"].append(p)
        out_cells[name] = c
    return {"library": lib.args[0] if lib.args else "", "units_si": units, "voltage_map": vmap,
            "nom": nom, "operating_conditions": pvt, "cells": out_cells,
            "units": {"nom_voltage": str(nom["voltage"])}}
