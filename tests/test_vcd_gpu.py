"""Device VCD ingest (gtp_vcd_parse) against the reading restatement
(oracle/vcd.cpp): every stream bit-exact (offsets, times, values), the
declarations, t_max and time scale; the same inputs rejected.  Then power
from a device-parsed VCD against power from the same streams loaded as
arrays, and against the original trace when every net is a scalar (GPU)."""
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.abi import GtpError
from gtpower_amd.power import HipPower
from gtpower_amd.vcd import VcdText

sys.path.insert(0, str(Path(__file__).resolve().parent))
import vcd_cases as V  # noqa: E402
from test_vcd_oracle import two_state_buses  # noqa: E402

pytestmark = pytest.mark.gpu

GCD = Path(__file__).resolve().parent / "golden" / "gcd_sky130hd.vcd"


@pytest.fixture(scope="module")
def reader():
    v = VcdText(device=0)
    yield v
    v.close()


def same_as_oracle(reader, text, what):
    print(f"[vcd] {what}: {len(text)} bytes", flush=True)
    info = reader.parse(text)
    ref = oracle.vcd_parse(text)
    assert [(n, i, w, f) for n, i, w, _, f in reader.vars()] == ref["vars"], what
    off, t, v = reader.events()
    assert np.array_equal(off, ref["offsets"]), what
    assert np.array_equal(t, ref["times"]), what
    assert np.array_equal(v, ref["values"]), what
    assert info["t_max"] == ref["t_max"] and info["time_scale"] == pytest.approx(ref["time_scale"], rel=1e-12)
    assert info["n_event"] == len(ref["times"]) and info["n_id"] == ref["n_id"]
    return info


@pytest.mark.parametrize("name,text,expect", V.CASES, ids=[c[0] for c in V.CASES])
def test_cases(reader, name, text, expect):
    same_as_oracle(reader, text, name)


@pytest.mark.parametrize("name,text", V.ERRORS, ids=[c[0] for c in V.ERRORS])
def test_rejected_like_the_oracle(reader, name, text):
    print(f"[vcd] error case {name}", flush=True)
    with pytest.raises(oracle.VcdError):
        oracle.vcd_parse(text)
    with pytest.raises(GtpError):
        reader.parse(text)
    same_as_oracle(reader, V.BASIC, "basic after an error")  # the reader stays usable


def test_gcd_dump(reader):
    info = same_as_oracle(reader, GCD.read_bytes(), "gcd_sky130hd.vcd")
    assert info["n_var"] == 7585 and info["t_max"] == 125000


def test_synthetic_with_buses(reader):
    d = S.make_design(3000, dff_fraction=0.15, clock_ns=2.0, seed=51)
    tr = d.trace(2000, toggle_p=0.2, seed=52)
    text = S.vcd_text(tr, buses=two_state_buses(tr, d.n_net))
    info = same_as_oracle(reader, text, "synthetic 3000 gates x 2000 cycles")
    assert info["n_stream"] == d.n_net


def test_text_edges(reader):
    # CRLF line ends, tabs, no newline at the end (the last token is dropped),
    # a text that ends right after the header
    same_as_oracle(reader, V.BASIC.replace(b"\n", b"\r\n"), "crlf")
    same_as_oracle(reader, V.BASIC.replace(b"\n", b"\t \n"), "tabs")
    same_as_oracle(reader, V.BASIC.rstrip(b"\n"), "no final newline")
    same_as_oracle(reader, V.QUIRKS + b"\n", "quirks with a final newline")
    # a repeat at the same time is dropped (the order check runs and passes);
    # the time goes back, but no id's kept times do
    same_as_oracle(reader, V._HEAD + b"1!\n#5\n0!\n0!\n#9\n", "a repeat at one time")
    same_as_oracle(reader, V._HEAD + b"1!\n#5\n0!\n#3\nb11 \"\n#9\n", "time back, per id forward")
    # digits past 32-bit arithmetic: a 40-digit and a saturating 70-digit bus
    # value, times of 9, 10 and 12 digits (the last two after the changes)
    same_as_oracle(reader, b"$timescale 1ps $end\n$var wire 64 ! w [63:0] $end\n$var wire 1 # a $end\n"
                           b"$enddefinitions $end\n#0\nb" + b"1" * 40 + b" !\n1#\n#7\nb" + b"10" * 35 +
                           b" !\n0#\n#123456789\nb101 !\n#1234567890\n#123456789012\n", "long digit strings")
    # bus values whose ids look like scalar changes ("x", "1a"), beside a
    # scalar change "1a" of id "a"
    same_as_oracle(reader, b"$timescale 1ps $end\n$var wire 2 x v [1:0] $end\n$var wire 2 1a w [1:0] $end\n"
                           b"$var wire 1 a s $end\n$enddefinitions $end\n#0\nb10 x\nb01 1a\n1a\n#5\nb11 x\n"
                           b"b11 1a\n0a\n#9\n", "bus ids that look like scalar changes")


@pytest.mark.parametrize("with_buses", [False, True])
def test_power_from_device_vcd(with_buses):
    """Power of a design over its trace read from VCD text on the device:
    identical to the same streams loaded as arrays (the oracle's reading),
    and, with scalar nets only, to the original trace."""
    d = S.make_design(2500, dff_fraction=0.15, clock_ns=2.0, seed=61)
    tr = d.trace(1500, toggle_p=0.2, seed=62)
    buses = two_state_buses(tr, d.n_net, max_width=16) if with_buses else []
    text = S.vcd_text(tr, buses=buses)
    ref = oracle.vcd_parse(text)
    arr = S.ArrayTrace(ref["offsets"], ref["times"], ref["values"], ref["t_max"], ref["time_scale"],
                       tr.clk_period, tr.ticks_per_cycle)
    print(f"[vcd] power from device vcd (buses={len(buses)})", flush=True)
    with VcdText(device=0) as v, HipPower() as e1, HipPower() as e2:
        v.parse(text)
        e1.set_design(d.library, d)
        e1.read_vcd(v, tr.clk_period)
        a = e1.power()
        e2.set_design(d.library, d)
        e2.read_power_activities(arr)
        b = e2.power()
        for k in ("gate_power", "cycle_power", "gate_toggle_x2", "gate_glitch"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
        if not buses:
            e2.read_power_activities(tr)
            c = e2.power()
            for k in ("gate_power", "cycle_power", "gate_toggle_x2", "gate_glitch"):
                assert np.array_equal(getattr(a, k), getattr(c, k)), k
