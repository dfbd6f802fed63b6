"""Cell library: Liberty-subset JSON -> the flat arrays of `gtp_library`.

The reference builds its power lookup structures from OpenSTA's Liberty objects
(power/TimeBasedPower.cc:258-301 getLeakagePower, :557-587 getStateIdx;
power/cuda/CudaPower.cu:2347-2624 getInput/OutputInternalPower).  Here the same
information is flattened once, on the host, into SI-unit arrays:

  * leakage: a 2^n_pin table per cell.  Every state starts at
    cell_leakage_power (or 0 when absent); each leakage_power group then
    overwrites the states its `when` matches, in Liberty order
    (TimeBasedPower.cc:289-299).  A `when` must be a conjunction of pin literals
    to match anything (PowerUtils.cc:110-153 collectWhenPinStates).
  * internal power: one gtp_internal_power record per Liberty internal_power
    group, with its `when` as a truth table over gate states (X counts as 0,
    PowerUtils.cc:155-187 evalWhen).
  * timing arcs (generator only): delay/slew tables per (output, input,
    from_rf, to_rf), from timing_sense / timing_type.
"""
import json
from pathlib import Path

import numpy as np

from . import abi

DATA_JSON = Path(__file__).resolve().parent / "data" / "sky130hd_tt_subset.json"

NS = 1e-9        # Liberty time unit (time_unit : "1ns")
PF = 1e-12       # capacitive_load_unit (1, pf)
NW = 1e-9        # leakage_power_unit : "1nW"
PJ = 1e-12       # internal energy unit: V^2 * pF


# --------------------------------------------------------------------------
# Boolean expressions of Liberty `function` / `when` strings.
# Grammar: expr := term ('|' | '+') term ... ; term := unary ('&' | '*') unary ...
#          xor  := '^' between terms ; unary := '!' unary | '(' expr ')' | name | 0 | 1
# --------------------------------------------------------------------------
def _lex(s):
    out, i = [], 0
    while i < len(s):
        c = s[i]
        if c.isspace():
            i += 1
        elif c in "!&|()^*+'":
            out.append(c)
            i += 1
        else:
            j = i
            while j < len(s) and (s[j].isalnum() or s[j] in "_[]"):
                j += 1
            if j == i:
                raise ValueError(f"bad character {c!r} in {s!r}")
            out.append(s[i:j])
            i = j
    return out


def parse_expr(s):
    """Parse into a tuple tree: ('port', name) | ('not', x) | ('and'|'or'|'xor', a, b) | ('one',) | ('zero',)."""
    toks = _lex(s)
    pos = [0]

    def peek():
        return toks[pos[0]] if pos[0] < len(toks) else None

    def take():
        t = toks[pos[0]]
        pos[0] += 1
        return t

    def unary():
        t = take()
        if t == "!":
            node = ("not", unary())
        elif t == "(":
            node = expr()
            if take() != ")":
                raise ValueError(f"missing ) in {s!r}")
        elif t == "1":
            node = ("one",)
        elif t == "0":
            node = ("zero",)
        else:
            node = ("port", t)
        while peek() == "'":
            take()
            node = ("not", node)
        return node

    def term():
        node = unary()
        while peek() in ("&", "*") or (peek() not in (None, "|", "+", ")", "^")):
            if peek() in ("&", "*"):
                take()
            node = ("and", node, unary())
        return node

    def xterm():
        node = term()
        while peek() == "^":
            take()
            node = ("xor", node, term())
        return node

    def expr():
        node = xterm()
        while peek() in ("|", "+"):
            take()
            node = ("or", node, xterm())
        return node

    tree = expr()
    if pos[0] != len(toks):
        raise ValueError(f"trailing tokens in {s!r}")
    return tree


def eval_expr(tree, value_of):
    op = tree[0]
    if op == "port":
        return value_of(tree[1])
    if op == "not":
        return not eval_expr(tree[1], value_of)
    if op == "and":
        return eval_expr(tree[1], value_of) and eval_expr(tree[2], value_of)
    if op == "or":
        return eval_expr(tree[1], value_of) or eval_expr(tree[2], value_of)
    if op == "xor":
        return eval_expr(tree[1], value_of) != eval_expr(tree[2], value_of)
    if op == "one":
        return True
    if op == "zero":
        return False
    raise ValueError(op)


def conjunction_literals(tree, negated=False):
    """Pin literals of a conjunction {name: 0|1}, or None when `tree` is not one.

    Restates PowerUtils.cc:110-153 collectWhenPinStates: NOT flips polarity,
    AND (or OR under negation, by De Morgan) combines, a contradictory literal or
    an XOR/OR makes the condition unusable."""
    op = tree[0]
    if op == "port":
        return {tree[1]: 0 if negated else 1}
    if op == "not":
        return conjunction_literals(tree[1], not negated)
    if op == "one":
        return {} if not negated else None
    if op == "zero":
        return {} if negated else None
    if (op == "and" and not negated) or (op == "or" and negated):
        a = conjunction_literals(tree[1], negated)
        b = conjunction_literals(tree[2], negated)
        if a is None or b is None:
            return None
        for k, v in b.items():
            if k in a and a[k] != v:
                return None
            a[k] = v
        return a
    return None


def when_truth(when, pin_names):
    """Truth table of a `when` string over gate states (bit p = pin p is 1).
    Cells of more than abi.MAX_WHEN_PINS pins take unconditional groups only:
    all 64 bits set (no sky130 cell of 7 or 8 pins conditions its internal
    power)."""
    n = len(pin_names)
    if not when:
        return (1 << (1 << min(n, abi.MAX_WHEN_PINS))) - 1
    if n > abi.MAX_WHEN_PINS:
        raise ValueError(f"`when` on internal power of a {n}-pin cell (at most {abi.MAX_WHEN_PINS} pins)")
    tree = parse_expr(when)
    idx = {name: k for k, name in enumerate(pin_names)}
    mask = 0
    for s in range(1 << n):
        if eval_expr(tree, lambda name: bool((s >> idx[name]) & 1)):
            mask |= 1 << s
    return mask


def leakage_states(when, pin_names):
    """States matched by a leakage `when` (TimeBasedPower.cc:557-587 getStateIdx)."""
    n = len(pin_names)
    if not when:
        return list(range(1 << n))
    lits = conjunction_literals(parse_expr(when))
    if lits is None:
        return []
    idx = {name: k for k, name in enumerate(pin_names)}
    base, free = 0, []
    for k in range(n):
        name = pin_names[k]
        if name in lits:
            if lits[name] == 1:
                base |= 1 << k
        else:
            free.append(k)
    states = [base]
    for k in free:
        states += [s | (1 << k) for s in states]
    return states


class CellLibrary:
    """Cells flattened for the engine and the generator: by default the eight
    sky130 cells of the bundled JSON (sky130 units: ns, pF, nW, pJ); or any
    library read by `gtpower_amd.liberty.read_liberty` (doc=...), with its own
    units, table axis order and internal-power scale factors."""

    def __init__(self, json_path=DATA_JSON, cells=None, doc=None):
        if doc is None:
            doc = json.loads(Path(json_path).read_text())
        self.doc = doc
        u = doc.get("units_si", {"time": NS, "capacitance": PF, "leakage_power": NW, "energy": PJ})
        self.unit = u
        names = cells or list(doc["cells"].keys())
        self.cell_names = names
        self.full_names = [doc["cells"][n]["name"] for n in names]
        self.voltage = float(doc["units"].get("nom_voltage", "1.8"))
        self._tables = []          # list of (n1, n2, var1, var2, axis1, axis2, values)
        self._data = []
        self.cells = []
        ipowers = []
        for ci, short in enumerate(names):
            c = doc["cells"][short]
            ins = [p for p in c["pins"] if p["direction"] in ("input", "inout")]
            outs = [p for p in c["pins"] if p["direction"] in ("output", "inout")]
            pins = ins + outs
            pin_names = [p["name"] for p in pins]
            n = len(pins)
            if n > abi.MAX_CELL_PINS:
                raise ValueError(f"{short}: {n} pins > {abi.MAX_CELL_PINS}")
            pidx = {p: k for k, p in enumerate(pin_names)}
            lu = u["leakage_power"]
            leak_default = c["cell_leakage_power"] * lu if np.isfinite(c["cell_leakage_power"]) else 0.0
            leak = None
            if c["leakage_power"]:
                leak = np.full(1 << n, leak_default)
                for lp in c["leakage_power"]:
                    for s in leakage_states(lp["when"], pin_names):
                        leak[s] = lp["value"] * lu
            # internal power at the operating conditions (liberty/Liberty.cc:280-316)
            pscale = float(c.get("internal_power_scale", 1.0))
            clock_mask = 0
            for k, p in enumerate(pins):
                if p["clock"]:
                    clock_mask |= 1 << k
            # internal power groups, Liberty order within each pin
            for k, p in enumerate(pins):
                for ip in p["internal_power"]:
                    rel = pidx[ip["related_pin"]] if ip["related_pin"] else -1
                    rise = self._add_power_table(ip["rise_power"], pscale)
                    fall = self._add_power_table(ip["fall_power"], pscale)
                    ipowers.append((ci, k, rel, rise, fall, when_truth(ip["when"], pin_names)))
            # timing arcs for the generator: (out_k, in_k, from_rf, to_rf) -> (delay table, slew table)
            arcs = {}
            for ok, p in enumerate(outs):
                for tm in p["timing"]:
                    i = pidx[tm["related_pin"]]
                    sense, ttype = tm["timing_sense"], tm["timing_type"]
                    pairs = []
                    if ttype == "rising_edge":
                        pairs = [(0, 0), (0, 1)]
                    elif ttype == "falling_edge":
                        pairs = [(1, 0), (1, 1)]
                    elif sense == "positive_unate":
                        pairs = [(0, 0), (1, 1)]
                    elif sense == "negative_unate":
                        pairs = [(0, 1), (1, 0)]
                    else:
                        pairs = [(0, 0), (0, 1), (1, 0), (1, 1)]
                    for fr, to in pairs:
                        dt = tm["cell_rise"] if to == 0 else tm["cell_fall"]
                        st = tm["rise_transition"] if to == 0 else tm["fall_transition"]
                        arcs[(ok, i, fr, to)] = (self._add_timing_table(dt), self._add_timing_table(st))
            func = None
            outp = outs[0]
            is_ff = "ff" in c
            if not is_ff:
                tree = parse_expr(outp["function"])
                in_names = [p["name"] for p in ins]
                func = 0
                for s in range(1 << len(ins)):
                    if eval_expr(tree, lambda nm: bool((s >> in_names.index(nm)) & 1)):
                        func |= 1 << s
            self.cells.append(dict(
                short=short, name=c["name"], pins=pin_names, n_pin=n, n_in=len(ins),
                n_out=len(outs), pin_cap=[p["capacitance"] * u["capacitance"] for p in pins],
                leak=leak, leak_default=leak_default, clock_mask=clock_mask,
                cls=abi.CLASS_SEQ if is_ff else abi.CLASS_COMB, arcs=arcs, func=func,
                is_ff=is_ff,
                # a plain pin name, else -1 (`!CLK_N`, `(D&!SCE) | (SCD&SCE)`)
                ff_clock=pidx.get(c["ff"]["clocked_on"], -1) if is_ff else -1,
                ff_data=pidx.get(c["ff"]["next_state"], -1) if is_ff else -1))
        self.ipowers = ipowers
        self._finish()

    # ---- table storage ------------------------------------------------
    def _add_table(self, t, var1, var2, yscale, xscale1, xscale2):
        if t is None:
            return -1
        if "vars" in t:
            # each axis named by its template variable, in its unit
            var1 = t["vars"][0]
            var2 = t["vars"][1] if len(t["vars"]) > 1 else abi.AXIS_CAP
            ax_unit = {abi.AXIS_SLEW: self.unit["time"], abi.AXIS_CAP: self.unit["capacitance"]}
            xscale1, xscale2 = ax_unit[var1], ax_unit[var2]
        ax1 = np.asarray(t["index_1"], dtype=np.float64) * xscale1
        vals = np.asarray(t["values"], dtype=np.float64) * yscale
        off1 = sum(len(d) for d in self._data)
        self._data.append(ax1)
        if "index_2" in t:
            ax2 = np.asarray(t["index_2"], dtype=np.float64) * xscale2
            off2 = off1 + len(ax1)
            self._data.append(ax2)
            n2 = len(ax2)
        else:
            off2, n2 = 0, 0
        offv = sum(len(d) for d in self._data)
        self._data.append(vals.reshape(-1))
        self._tables.append((len(ax1), n2, var1, var2, off1, off2, offv))
        return len(self._tables) - 1

    def _add_power_table(self, t, scale=1.0):
        # sky130 power_inputs_1: input_transition_time; power_outputs_1:
        # (input_transition_time, total_output_net_capacitance); energies in
        # voltage_unit^2 * capacitive_load_unit
        u = self.unit
        return self._add_table(t, abi.AXIS_SLEW, abi.AXIS_CAP, u["energy"] * scale, u["time"], u["capacitance"])

    def _add_timing_table(self, t):
        # sky130 del_1_7_7: input_net_transition x total_output_net_capacitance, values in time_unit
        u = self.unit
        return self._add_table(t, abi.AXIS_SLEW, abi.AXIS_CAP, u["time"], u["time"], u["capacitance"])

    def _finish(self):
        self.table_data = np.concatenate(self._data) if self._data else np.zeros(1)
        self.tables = np.array(self._tables, dtype=np.int64).reshape(-1, 7)
        nc = len(self.cells)
        self.cell_n_pin = np.array([c["n_pin"] for c in self.cells], dtype=np.int32)
        self.cell_n_in = np.array([c["n_in"] for c in self.cells], dtype=np.int32)
        self.cell_n_out = self.cell_n_pin - self.cell_n_in
        self.cell_class = np.array([c["cls"] for c in self.cells], dtype=np.int32)
        self.cell_clock_mask = np.array([c["clock_mask"] for c in self.cells], dtype=np.uint32)
        offs = [0]
        vals = []
        for c in self.cells:
            if c["leak"] is not None:
                vals.append(c["leak"])
                offs.append(offs[-1] + len(c["leak"]))
            else:
                offs.append(offs[-1])
        self.cell_leak_offset = np.array(offs, dtype=np.int64)
        self.leak_values = np.concatenate(vals) if vals else np.zeros(1)
        self.cell_leak_default = np.array([c["leak_default"] for c in self.cells], dtype=np.float64)
        ip = np.zeros(len(self.ipowers), dtype=[("cell", "<i4"), ("pin", "<i4"), ("related_pin", "<i4"),
                                                ("rise_table", "<i4"), ("fall_table", "<i4"),
                                                ("pad_", "<i4"), ("when_truth", "<u8")])
        for k, (ci, pk, rel, r, f, w) in enumerate(self.ipowers):
            ip[k] = (ci, pk, rel, r, f, 0, w)
        self.ipower_array = ip
        tb = np.zeros(len(self.tables), dtype=[("n1", "<i4"), ("n2", "<i4"), ("var1", "<i4"),
                                               ("var2", "<i4"), ("axis1", "<i8"), ("axis2", "<i8"),
                                               ("values", "<i8")])
        for k, row in enumerate(self.tables):
            tb[k] = tuple(int(x) for x in row)
        self.table_array = tb
        assert nc == len(self.cell_names)

    def index(self, short):
        return self.cell_names.index(short)

    # ---- ABI ------------------------------------------------------------
    def as_abi(self):
        """A GtpLibrary pointing into this object's arrays (keep `self` alive)."""
        lib = abi.GtpLibrary()
        lib.n_cell = len(self.cells)
        lib.cell_n_pin = abi.ptr(self.cell_n_pin, abi.C.c_int32)
        lib.cell_n_in = abi.ptr(self.cell_n_in, abi.C.c_int32)
        lib.cell_class = abi.ptr(self.cell_class, abi.C.c_int32)
        lib.cell_clock_mask = abi.ptr(self.cell_clock_mask, abi.C.c_uint32)
        lib.cell_leak_offset = abi.ptr(self.cell_leak_offset, abi.C.c_int64)
        lib.leak_values = abi.ptr(self.leak_values, abi.C.c_double)
        lib.cell_leak_default = abi.ptr(self.cell_leak_default, abi.C.c_double)
        lib.n_table = len(self.table_array)
        lib.tables = self.table_array.ctypes.data_as(abi.C.POINTER(abi.GtpTable))
        lib.table_data = abi.ptr(self.table_data, abi.C.c_double)
        lib.n_ipower = len(self.ipower_array)
        lib.ipowers = self.ipower_array.ctypes.data_as(abi.C.POINTER(abi.GtpInternalPower))
        return lib

    # ---- generator view ---------------------------------------------------
    def arc_table_index(self):
        """int32 [n_cell, MAX_OUT=1, MAX_PIN, 2, 2, 2]: (delay, slew) table per arc, -1 = none."""
        nc = len(self.cells)
        out = np.full((nc, abi.MAX_CELL_PINS, 2, 2, 2), -1, dtype=np.int32)
        for ci, c in enumerate(self.cells):
            for (ok, i, fr, to), (dt, st) in c["arcs"].items():
                if ok != 0:
                    raise ValueError("generator supports single-output cells")
                out[ci, i, fr, to, 0] = dt
                out[ci, i, fr, to, 1] = st
        return out

    def interp(self, table_index, x1, x2=0.0):
        """FP64 Liberty interpolation (liberty/TableModel.cc:858-1062) -- host helper."""
        n1, n2, _, _, a1, a2, v = (int(x) for x in self.tables[table_index])
        d = self.table_data
        ax1 = d[a1:a1 + n1]
        if n2 == 0:
            y = d[v:v + n1]
            return _interp1(ax1, y, x1)
        ax2 = d[a2:a2 + n2]
        y = d[v:v + n1 * n2].reshape(n1, n2)
        return _interp2(ax1, ax2, y, x1, x2)


def find_axis_index(axis, x):
    """liberty/TableModel.cc:1465-1488 findValueIndex."""
    n = len(axis)
    if n <= 1 or x <= axis[0]:
        return 0
    if x >= axis[-1]:
        return n - 2
    lo, hi = -1, n
    while hi - lo > 1:
        mid = (hi + lo) >> 1
        if x >= axis[mid]:
            lo = mid
        else:
            hi = mid
    return lo


def _interp1(ax, y, x):
    if len(ax) == 1:
        return float(y[0])
    i = find_axis_index(ax, x)
    dx = (x - ax[i]) / (ax[i + 1] - ax[i])
    return (1 - dx) * y[i] + dx * y[i + 1]


def _interp2(ax1, ax2, y, x1, x2):
    n1, n2 = len(ax1), len(ax2)
    if n1 == 1 and n2 == 1:
        return float(y[0, 0])
    if n1 == 1:
        return _interp1(ax2, y[0], x2)
    if n2 == 1:
        return _interp1(ax1, y[:, 0], x1)
    i = find_axis_index(ax1, x1)
    j = find_axis_index(ax2, x2)
    dx1 = (x1 - ax1[i]) / (ax1[i + 1] - ax1[i])
    dx2 = (x2 - ax2[j]) / (ax2[j + 1] - ax2[j])
    return ((1 - dx1) * (1 - dx2) * y[i, j] + dx1 * (1 - dx2) * y[i + 1, j]
            + dx1 * dx2 * y[i + 1, j + 1] + (1 - dx1) * dx2 * y[i, j + 1])
