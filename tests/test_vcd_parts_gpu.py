"""Device VCD ingest of a text in parts (GPU).

A text longer than the part size (1 GiB by default, GTP_VCD_PART_BYTES here)
is parsed in parts cut at whitespace: the time in force, every id's last
kept change and the error positions are carried from part to part, and the
parts' stores are appended per stream.  Small part sizes cut the
hand-worked cases, the reference's GCD testbench dump and a synthetic dump
with buses in many places: every stream must equal the oracle's reading of
the whole text bit for bit, the counts must equal those of the one-part
parse, and the errors must be the same (kind and item)."""
import sys
from pathlib import Path

import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.abi import GtpError
from gtpower_amd.vcd import VcdText

sys.path.insert(0, str(Path(__file__).resolve().parent))
import vcd_cases as V  # noqa: E402
from test_vcd_gpu import GCD, same_as_oracle  # noqa: E402
from test_vcd_oracle import two_state_buses  # noqa: E402

pytestmark = pytest.mark.gpu

COUNTS = ("n_var", "n_id", "n_stream", "n_event", "t_max", "n_token", "n_change", "body_bytes")


@pytest.fixture(scope="module")
def reader():
    v = VcdText(device=0)
    yield v
    v.close()


def in_parts(monkeypatch, reader, text, size, what):
    monkeypatch.delenv("GTP_VCD_PART_BYTES", raising=False)
    one = dict(reader.parse(text))
    assert one["n_part"] == 1
    monkeypatch.setenv("GTP_VCD_PART_BYTES", str(size))
    info = same_as_oracle(reader, text, f"{what} in parts of {size} bytes")
    for k in COUNTS:
        assert info[k] == one[k], (what, size, k)
    return info


TEXTS = [(n, t) for n, t, _ in V.CASES] + [
    ("quirks_newline", V.QUIRKS + b"\n"),
    ("basic_crlf", V.BASIC.replace(b"\n", b"\r\n")),
    # a bus value's id split from it by a newline, runs of tokens that look
    # like bus values (ids "b" and "bb"), changes before the first time token
    ("bus_runs", b"$timescale 1ps $end\n$var wire 2 b x [1:0] $end\n$var wire 1 bb y $end\n"
                 b"$var wire 1 ! z $end\n$enddefinitions $end\n$dumpvars\n1!\nb10\nb\nb1 bb\n#3\n"
                 b"b11 b b0 bb 0! #4 b11 b\nb1 bb 1! #9 b01 b b0 bb #12 b01 b 1! #13\n"),
]


@pytest.mark.parametrize("size", [1, 4, 16, 37, 100])
@pytest.mark.parametrize("name,text", TEXTS, ids=[t[0] for t in TEXTS])
def test_cases_in_parts(monkeypatch, reader, name, text, size):
    info = in_parts(monkeypatch, reader, text, size, name)
    if len(text) > 2 * size + 120:
        assert info["n_part"] > 1, (name, size)


@pytest.mark.parametrize("size", [3, 11, 40])
@pytest.mark.parametrize("name,text", V.ERRORS, ids=[c[0] for c in V.ERRORS])
def test_errors_in_parts(monkeypatch, reader, name, text, size):
    """The same error as the one-part parse; the item (the token or change it
    names, counted from the start of the value changes) is the same too, but
    for an order error, whose item counts changes grouped by id per part."""
    print(f"[vcd parts] error case {name} in parts of {size} bytes", flush=True)
    monkeypatch.delenv("GTP_VCD_PART_BYTES", raising=False)
    with pytest.raises(GtpError) as one:
        reader.parse(text)
    monkeypatch.setenv("GTP_VCD_PART_BYTES", str(size))
    with pytest.raises(GtpError) as parts:
        reader.parse(text)
    a, b = str(one.value), str(parts.value)
    if "not increasing" in a:
        assert a.split(" (item")[0] == b.split(" (item")[0]
    else:
        assert a == b
    same_as_oracle(reader, V.BASIC, "basic after an error")


@pytest.mark.parametrize("size", [4096, 50_000])
def test_gcd_dump_in_parts(monkeypatch, reader, size):
    info = in_parts(monkeypatch, reader, GCD.read_bytes(), size, "gcd_sky130hd.vcd")
    assert info["n_part"] >= info["body_bytes"] // size  # the declarations are in the first part


def test_synthetic_with_buses_in_parts(monkeypatch, reader):
    d = S.make_design(3000, dff_fraction=0.15, clock_ns=2.0, seed=51)
    tr = d.trace(2000, toggle_p=0.2, seed=52)
    text = S.vcd_text(tr, buses=two_state_buses(tr, d.n_net))
    size = len(text) // 23 + 1
    info = in_parts(monkeypatch, reader, text, size, "synthetic 3000 gates x 2000 cycles")
    assert info["n_part"] >= 23
