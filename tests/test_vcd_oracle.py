"""The VCD reading restatement (oracle/vcd.cpp) against hand-worked cases,
the reference's GCD testbench dump (tests/golden/gcd_sky130hd.vcd, from the
reference's test/gcd/data) read line by line independently, and a synthetic
round trip."""
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S

sys.path.insert(0, str(Path(__file__).resolve().parent))
import vcd_cases as V  # noqa: E402

GCD = Path(__file__).resolve().parent / "golden" / "gcd_sky130hd.vcd"


@pytest.mark.parametrize("name,text,expect", V.CASES, ids=[c[0] for c in V.CASES])
def test_known_answers(name, text, expect):
    r = oracle.vcd_parse(text)
    assert r["vars"] == expect["vars"]
    off, t, v = V.expected_arrays(expect)
    assert np.array_equal(r["offsets"], off)
    assert np.array_equal(r["times"], t)
    assert np.array_equal(r["values"], v)
    assert r["t_max"] == expect["t_max"]
    assert r["time_scale"] == pytest.approx(expect["time_scale"], rel=1e-12)


@pytest.mark.parametrize("name,text", V.ERRORS, ids=[c[0] for c in V.ERRORS])
def test_errors(name, text):
    with pytest.raises(oracle.VcdError):
        oracle.vcd_parse(text)


def _line_reader(path, ids):
    """Independent reading of a one-token-per-line dump: changes per id after
    the line "$dumpall" (the value section of this file), deduplicated."""
    out = {i: [] for i in ids}
    t = 0
    started = False
    for line in open(path, "rb").read().decode().splitlines():
        s = line.strip()
        if not started:
            started = s == "$dumpall"
            continue
        if s.startswith("#"):
            t = int(s[1:])
        elif s[:1] in ("0", "1", "x", "z") and s[1:] in out:
            out[s[1:]].append((t, {"0": 0, "1": 1, "x": 2, "z": 3}[s[0]], s[0]))
        elif s[:1] == "b" and " " in s and s.split()[1] in out:
            bits, i = s[1:].split()
            out[i].append((t, 0 if bits[0] in "xz" else int(bits, 2), bits))
    return out


def test_gcd_dump():
    text = GCD.read_bytes()
    r = oracle.vcd_parse(text)
    n_var_lines = sum(1 for ln in text.decode().splitlines() if ln.startswith("$var"))
    assert len(r["vars"]) == n_var_lines == 7585
    assert r["t_max"] == 125000 and r["time_scale"] == pytest.approx(1e-12)
    by_id = {}
    for name, i, w, first in r["vars"]:
        by_id.setdefault(i, (w, first))
    ref = _line_reader(GCD, [")", '"', "*", "P_"])
    off = r["offsets"]
    # clk: scalar, every 2500 ps, starting 0 at time 0
    w, s = by_id[")"]
    want = [(t, v) for t, v, _ in ref[")"]]
    want = [x for k, x in enumerate(want) if k == 0 or x[1] != want[k - 1][1]]
    assert list(zip(r["times"][off[s]:off[s + 1]], r["values"][off[s]:off[s + 1]])) == want
    assert len(want) == 51 and want[1] == (2500, 1)
    # resp_msg: a 16-bit bus, bit b of each kept value
    w, s = by_id['"']
    assert w == 16
    vals = []
    for t, v, _ in ref['"']:
        if not vals or vals[-1][1] != v:
            vals.append((t, v))
    for b in range(16):
        q = s + b
        assert list(r["times"][off[q]:off[q + 1]]) == [t for t, _ in vals]
        assert list(r["values"][off[q]:off[q + 1]]) == [(v >> b) & 1 for _, v in vals]
    # an X at the start of the dump
    w, s = by_id["P_"]
    assert r["values"][off[s]] == 2 and r["times"][off[s]] == 0


def two_state_buses(trace, n_net, max_width=64):
    """Bus ranges over runs of consecutive nets that never carry X or Z (a
    bus value with an X letter does not keep its 0/1 bits in the reference's
    reading)."""
    off = trace.net_event_offset
    two = np.array([np.all(trace.event_value[off[n]:off[n + 1]] < 2) for n in range(n_net)])
    buses, n = [], 0
    while n < n_net:
        if not two[n]:
            n += 1
            continue
        m = n
        while m < n_net and two[m] and m - n < max_width:
            m += 1
        if m - n >= 2:
            buses.append((n, m - n))
        n = m
    return buses


def test_synthetic_round_trip():
    d = S.make_design(400, dff_fraction=0.15, clock_ns=2.0, seed=41)
    tr = d.trace(300, toggle_p=0.2, seed=42)
    buses = two_state_buses(tr, d.n_net)
    assert len(buses) >= 3 and max(w for _, w in buses) >= 8
    r = oracle.vcd_parse(S.vcd_text(tr, buses=buses))
    assert r["time_scale"] == pytest.approx(tr.time_scale) and r["t_max"] == tr.t_max
    assert len(r["offsets"]) - 1 == d.n_net
    in_bus = np.zeros(d.n_net, bool)
    for f, w in buses:
        in_bus[f:f + w] = True
    off, ro = tr.net_event_offset, r["offsets"]
    for n in range(d.n_net):
        a_t, a_v = tr.event_time[off[n]:off[n + 1]], tr.event_value[off[n]:off[n + 1]]
        b_t, b_v = r["times"][ro[n]:ro[n + 1]], r["values"][ro[n]:ro[n + 1]]
        if not in_bus[n]:
            assert np.array_equal(a_t, b_t) and np.array_equal(a_v, b_v)
        else:
            # a bus bit is written whenever any bit of its bus changes: its own
            # changes are the value changes of its stream
            keep = np.r_[True, b_v[1:] != b_v[:-1]]
            assert np.array_equal(b_t[keep], a_t) and np.array_equal(b_v[keep], a_v)
