"""Writes an excerpt of the reference's sky130 tt library: the library header
(everything before its first cell group), the named cells in the order given,
and the closing brace.  Run in a checkout that holds the reference:

    python tests/golden/make_sky130_excerpt.py OUT.lib CELL [CELL ...]

The fixtures it made:
    sky130hd_tt_8cells.lib       inv_1 buf_1 nand2_1 nor2_1 xor2_1 a21oi_1 mux2_1 dfxtp_1
    sky130hd_tt_large_cells.lib  mux4_1 sdfbbp_1 sdfrbp_1 a222oi_1
(cell names without the sky130_fd_sc_hd__ prefix)."""
import re
import sys
from pathlib import Path

SRC = Path("/root/reference/examples/sky130hd_tt.lib")
PREFIX = "sky130_fd_sc_hd__"


def cell_blocks(text):
    """name -> the cell group's text, from its `cell (` line to its closing brace."""
    out = {}
    for m in re.finditer(r'^    cell \("([^"]+)"\) \{\n', text, re.M):
        depth, i = 1, m.end()
        while depth:
            c = text[i]
            depth += c == "{"
            depth -= c == "}"
            i += 1
        out[m.group(1)] = text[m.start():i] + "\n"
    return out


def main():
    out, names = Path(sys.argv[1]), sys.argv[2:]
    text = SRC.read_text()
    first = re.search(r'^    cell \(', text, re.M).start()
    blocks = cell_blocks(text)
    body = "".join(blocks[PREFIX + n] for n in names)
    out.write_text(text[:first] + body + "}\n")


if __name__ == "__main__":
    main()
