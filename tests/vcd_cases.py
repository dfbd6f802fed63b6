"""Hand-made VCD texts with their expected per-(id, bit) streams, worked out
by hand from the reference's reading rules (power/VcdReader.cc:80-131,
:238-310; power/Vcd.cc:165-232; value table power/cuda/CudaPower.cu:1129-1145,
U read as X).  Shared by the oracle tests and the device ingest tests."""

BASIC = b"""$timescale 1ns $end
$var wire 1 ! a $end
$var wire 3 " bus [2:0] $end
$enddefinitions $end
#0
$dumpvars
x!
b000 "
$end
#5
1!
b101 "
#7
1!
b101 "
#9
0!
b11 "
#12
z!
bx "
#15
"""
BASIC_EXPECT = dict(
    vars=[("a", "!", 1, 0), ("bus[2:0]", '"', 3, 1)],
    streams=[([0, 5, 9, 12], [2, 1, 0, 3]),
             ([0, 5, 9, 12], [0, 1, 1, 0]),   # bit 0 of 000, 101, 011, 0 (x)
             ([0, 5, 9, 12], [0, 0, 1, 0]),   # bit 1
             ([0, 5, 9, 12], [0, 1, 0, 0])],  # bit 2
    t_max=15, time_scale=1e-9)

# The first token after the declarations starts the value changes and is not
# one of them (so "#100" is lost and times stay 0); "#x" and "b" are ids; id
# "b" gets a scalar, then a bus, then a scalar; "!" is declared twice; "%" is
# a real (its values are skipped); a comment's words are read as tokens
# ("1!" among them); the last token, cut by the end of the text, is dropped.
QUIRKS = (b"""$timescale 10 ps $end
$scope module m $end
$var reg 1 b a $end
$var wire 2 #x c [1:0] $end
$var real 1 % r $end
$var wire 1 ! d $end
$var wire 1 ! d2 $end
$upscope $end
$enddefinitions $end
#100
1b
b10 #x
r3.5 %
U!
#20
b1 b
$comment hello 1! world $end
#30
0b
b10 #x
u!
#40
x!
#50
b11 #x
#60
#70""")
QUIRKS_EXPECT = dict(
    vars=[("m/a", "b", 1, 0), ("m/c[1:0]", "#x", 2, 1), ("m/r", "%", 1, 3), ("m/d", "!", 1, 4),
          ("m/d2", "!", 1, 4)],
    streams=[([0, 20, 30], [1, 1, 0]),       # '1', bus 1 (bit 0), '0'
             ([0, 50], [0, 1]),              # bus 2, bus 3: bit 0
             ([0, 50], [1, 1]),              # bit 1
             ([], []),                       # the real
             ([0, 20, 30, 40], [2, 1, 2, 2])],  # U, 1, U, X
    t_max=60, time_scale=1e-11)

# first changes after time 0: X (scalar) or 0 (bus) is inserted at 0;
# upper-case value letters; an undeclared-width bus string shorter than the
# declaration; the text ends with whitespace
LATE = b"""$timescale 1ps $end
$var wire 1 ! a $end
$var wire 4 " v [3:0] $end
$var wire 1 # q $end
$enddefinitions $end
$dumpvars
#10
Z!
B11 "
#20
X!
b0011 "
1#
#25
0#
"""
LATE_EXPECT = dict(
    vars=[("a", "!", 1, 0), ("v[3:0]", '"', 4, 1), ("q", "#", 1, 5)],
    streams=[([0, 10, 20], [2, 3, 2]),
             ([0, 10], [0, 1]), ([0, 10], [0, 1]), ([0, 10], [0, 0]), ([0, 10], [0, 0]),
             ([0, 20, 25], [2, 1, 0])],
    t_max=25, time_scale=1e-12)

HEADER_ONLY = b"""$timescale 1ns $end
$var wire 1 ! a $end
$enddefinitions $end
"""
HEADER_ONLY_EXPECT = dict(vars=[("a", "!", 1, 0)], streams=[([], [])], t_max=0, time_scale=1e-9)

CASES = [("basic", BASIC, BASIC_EXPECT), ("quirks", QUIRKS, QUIRKS_EXPECT), ("late", LATE, LATE_EXPECT),
         ("header_only", HEADER_ONLY, HEADER_ONLY_EXPECT)]

_HEAD = b"$timescale 1ps $end\n$var wire 1 ! a $end\n$var wire 2 \" b [1:0] $end\n$enddefinitions $end\n#0\n"
ERRORS = [
    ("unknown_id", _HEAD + b"1!\n#5\n1?\n#6\n"),
    ("time_backwards", _HEAD + b"1!\n#5\n0!\n#3\n1!\n#9\n"),
    ("same_time", _HEAD + b"1!\n#5\n0!\n1!\n#9\n"),  # two kept changes of ! at 5
    ("time_range", _HEAD + b"1!\n#600000000\n0!\n#600000001\n"),
    ("time_syntax", _HEAD + b"1!\n#abc\n0!\n#9\n"),
    ("bus_without_id", _HEAD + b"1!\n#5\nb10\n"),
    ("unknown_unit", b"$timescale 1us $end\n$var wire 1 ! a $end\n$enddefinitions $end\n#0\n1!\n#5\n"),
    ("unhandled_command", b"$timescale 1ps $end\n$bogus $end\n$var wire 1 ! a $end\n#0\n1!\n#5\n"),
]


def expected_arrays(expect):
    """(offsets, times, values) of an expectation."""
    import numpy as np
    off = [0]
    times, vals = [], []
    for t, v in expect["streams"]:
        times += t
        vals += v
        off.append(off[-1] + len(t))
    return np.array(off, np.int64), np.array(times, np.int64), np.array(vals, np.uint8)
