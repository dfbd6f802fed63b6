"""Cells of 7 and 8 pins on the GPU: sky130 mux4_1, a222oi_1 and the scan
flops with set/reset (sdfbbp_1, sdfrbp_1), read from the reference's sky130
tt data with gtpower_amd.liberty, beside the small cells, on random netlists
and waveforms (X values, unconnected and silent inputs, absent arcs).  Same
bar as tests/test_engine_gpu.py: counts bit-exact, every power value within
1e-6 x the oracle's sum of |term| magnitudes.  The reference states leakage
and internal power by pin state up to 16 pins (util/GlobalConfig.hh:42-43,
power/TimeBasedPower.cc:258-301)."""
import numpy as np
import pytest

import oracle
from gtpower_amd import abi
from gtpower_amd.power import HipPower

from hand_libs import random_hand_design, random_hand_trace, sky130_library
from test_engine_gpu import compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("large_share,tile,seed", [(0.5, 8, 1), (1.0, 1, 2), (0.3, 128, 3)])
def test_large_cells_match_the_oracle(large_share, tile, seed):
    print(f"[large] 400 gates, large share {large_share}, tile {tile}, seed {seed}", flush=True)
    lib = sky130_library()
    d = random_hand_design(lib, 400, 40, seed, large_share=large_share)
    tr = random_hand_trace(d.n_net, 300, seed + 50)
    with HipPower(tile_cycles=tile) as eng:
        eng.set_design(lib, d)
        eng.read_power_activities(tr)
        res = eng.power()
    ref = oracle.run(lib, d, tr, threads=0)
    compare(res, ref, f"large cells {large_share}/{tile}/{seed}")
    big = np.isin(d.gate_cell, [lib.index(c) for c in ("mux4_1", "sdfbbp_1", "sdfrbp_1", "a222oi_1")])
    assert big.sum() > 0 and res.gate_power[:, big].sum() > 0 and ref["gate_glitch"][big].sum() > 0


def test_large_cells_sliced_and_batched_equal_the_resident_run():
    print("[large] sliced and batched runs vs the resident run", flush=True)
    lib = sky130_library()
    d = random_hand_design(lib, 300, 30, 7)
    trs = [random_hand_trace(d.n_net, n, 70 + n) for n in (500, 90)]
    with HipPower(tile_cycles=16) as eng:
        eng.set_design(lib, d)
        singles = []
        for tr in trs:
            eng.read_power_activities(tr)
            singles.append(eng.power())
        sliced = eng.power_sliced(trs[0], max_events=trs[0].n_event // 3 + 4 * d.n_net)
        assert sliced.schedule["n_interval"] >= 2
        batch = eng.power_batch(trs)
    for a, b in [(singles[0], sliced)] + list(zip(singles, batch)):
        assert np.array_equal(a.gate_toggle_x2, b.gate_toggle_x2)
        assert np.array_equal(a.gate_glitch, b.gate_glitch)
        assert np.array_equal(a.gate_power, b.gate_power)
        assert np.array_equal(a.cycle_power, b.cycle_power)
    compare(singles[1], oracle.run(lib, d, trs[1], threads=0), "large cells, short trace")


def test_when_on_internal_power_of_a_large_cell_is_refused():
    """Internal power `when` tables hold 64 states: a conditioned group on a
    cell of 7 or more pins is GTP_ERR_UNSUPPORTED, and the context stays usable."""
    print("[large] conditioned internal power on a 7-pin cell", flush=True)
    lib = sky130_library(large_only=True)
    d = random_hand_design(lib, 20, 8, 11)
    tr = random_hand_trace(d.n_net, 40, 12)
    ip = lib.ipower_array
    k = int(np.flatnonzero(ip["cell"] == lib.index("mux4_1"))[0])
    saved = ip["when_truth"][k]
    ip["when_truth"][k] = np.uint64(0x5555555555555555)
    with HipPower() as eng:
        with pytest.raises(abi.GtpError) as e:
            eng.set_design(lib, d)
        assert e.value.status == -5
        ip["when_truth"][k] = saved
        eng.set_design(lib, d)
        eng.read_power_activities(tr)
        res = eng.power()
    compare(res, oracle.run(lib, d, tr, threads=0), "after the refusal")


@pytest.mark.parametrize("tile", [4, 64])
def test_cells_of_9_12_and_16_pins_match_the_oracle(tile):
    """Hand-written cells of 9, 12 and 16 pins (the 16-pin kernel variant
    with unused pin slots and without; 2^16-state leakage tables read from
    device memory) beside the sky130 cells, two evaluations in a row."""
    from hand_libs import wide_library
    print(f"[large] 9-, 12- and 16-pin cells, tile {tile}", flush=True)
    lib = wide_library(sizes=(8, 11, 15))
    wide_cells = ("WIDE9", "WIDE12", "WIDE16")
    d = random_hand_design(lib, 300, 40, 21, large_share=0.5, large_cells=wide_cells + ("sdfbbp_1",))
    tr = random_hand_trace(d.n_net, 200, 22)
    with HipPower(tile_cycles=tile) as eng:
        eng.set_design(lib, d)
        eng.read_power_activities(tr)
        res = [eng.power() for _ in range(2)]
    ref = oracle.run(lib, d, tr, threads=0)
    for k, r in enumerate(res):
        compare(r, ref, f"wide cells, tile {tile}, evaluation {k}")
    for c in wide_cells:
        wide = d.gate_cell == lib.index(c)
        assert wide.sum() > 10 and res[0].gate_power[:, wide].sum() > 0 and ref["gate_glitch"][wide].sum() > 0


def test_many_cell_types_and_large_tables():
    """72 cell types and over 12 000 table values: the cell records and the
    Liberty tables no longer fit the kernels' LDS staging and are read in
    place."""
    from hand_libs import wide_library
    print("[large] 72 cell types", flush=True)
    lib = wide_library(sizes=(11, 15) + (1, 2, 3) * 20)
    assert len(lib.cells) == 74 and len(lib.table_data) > 12000
    d = random_hand_design(lib, 600, 40, 31, large_share=0.2)
    tr = random_hand_trace(d.n_net, 150, 32)
    with HipPower(tile_cycles=32) as eng:
        eng.set_design(lib, d)
        eng.read_power_activities(tr)
        res = eng.power()
    compare(res, oracle.run(lib, d, tr, threads=0), "72 cell types")
    assert len(np.unique(d.gate_cell)) > 64


def test_leakage_precision_beside_a_cell_of_huge_leakage():
    """A cell leaking up to 3.5e5 nW (as sky130's lpflow_bleeder_1 does,
    1e5 nW) beside sky130 cells whose states leak 1e-5 nW: every gate's
    leakage still meets the bar against its own magnitude (per-gate sums count
    each cell's own quanta)."""
    from gtpower_amd import liberty as L
    from gtpower_amd.library import CellLibrary
    from hand_libs import wide_cell_liberty
    print("[large] leakage precision beside a 3.5e5 nW cell", flush=True)
    lib = sky130_library()
    doc = lib.doc
    doc["cells"].update(L.read_liberty(wide_cell_liberty(2, seed=5, name="BLEED", leak_scale=1e5))["cells"])
    lib = CellLibrary(doc=doc)
    d = random_hand_design(lib, 400, 40, 41, large_share=0.02, large_cells=("BLEED",))
    tr = random_hand_trace(d.n_net, 150, 42)
    bleed = d.gate_cell == lib.index("BLEED")
    assert 1 <= bleed.sum() <= 30
    with HipPower(tile_cycles=16) as eng:
        eng.set_design(lib, d)
        eng.read_power_activities(tr)
        res = eng.power()
    ref = oracle.run(lib, d, tr, threads=0)
    compare(res, ref, "bleeder")
    small = ref["gate_power"][4, ~bleed]
    assert small.min() < 1e-7 * ref["gate_power"][4, bleed].max()
