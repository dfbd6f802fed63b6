"""GPU parity cases beyond the random designs: the hand-worked `when` and
related-input cases, a precision probe of the fixed-point accumulation, the
full-size small workload, and the reference's two output files written from a
GPU run.  Same bar as tests/test_engine_gpu.py: counts bit-exact, every power
value within 1e-6 x the oracle's sum of |term| magnitudes."""
import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.library import CellLibrary
from gtpower_amd.power import HipPower

from hand_libs import WHEN_CASES, when_case
from test_engine_gpu import compare, run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", [c[0] for c in WHEN_CASES])
@pytest.mark.parametrize("tile", [1, 4])
def test_when_and_related_cases(name, tile):
    print(f"[parity] when case {name} tile {tile}", flush=True)
    lib, d, tr = when_case(name)
    res, ref = run_both(d, tr, tile_cycles=tile)
    compare(res, ref, name)


def probe_design(lib, big_cap=5e-12, small_cap=0.2e-15, small_slew=5e-12):
    """A buf_1 driving `big_cap` (the largest energy term of the design: its
    switching energy) beside an inv_1 with a tiny load and fast input slew
    (energy terms below 2e-4 of the largest)."""
    cb, ci = lib.index("buf_1"), lib.index("inv_1")
    slew = [50e-12, 50e-12, 80e-12, 80e-12, small_slew, small_slew, 8e-12, 8e-12]
    arcs = np.full(8, np.nan)
    arcs[0 * 4 + 0] = 50e-12   # buf A rise -> X rise
    arcs[0 * 4 + 3] = 55e-12   # buf A fall -> X fall
    arcs[4 + 1] = 12e-12       # inv A rise -> Y fall
    arcs[4 + 2] = 14e-12       # inv A fall -> Y rise
    return S.ArrayDesign(lib, [cb, ci], [0, 1, 2, 3], slew, [0.0, big_cap, 0.0, small_cap], arcs, 4)


def probe_trace(n_cycle=64, T=1000):
    """The buf toggles in the first half of the cycles only; the inv in every
    cycle (plus a glitch pair every 8th cycle), so the second half's per-cycle
    power is the inv's alone."""
    w = [[(0, 0)], [(0, 0)], [(0, 0)], [(0, 1)]]
    vb = vi = 0
    for c in range(n_cycle):
        t0 = c * T
        if c < n_cycle // 2:
            vb ^= 1
            w[0].append((t0 + 100, vb))
            w[1].append((t0 + 150 + 5 * (vb == 0), vb))
        vi ^= 1
        w[2].append((t0 + 200, vi))
        w[3].append((t0 + 213, 1 - vi))
        if c % 8 == 3:
            vi ^= 1
            w[2].append((t0 + 400, vi))
            w[3].append((t0 + 409, 1 - vi))
    return S.trace_from_lists(w, n_cycle * T, 1e-12, T * 1e-12)


def test_small_gate_beside_the_largest_load():
    """The fixed-point quantum is set by the largest energy term of the design;
    a gate whose terms are below 2e-4 of it must still meet the bar per gate and
    per cycle (the cycles where it is the only active gate included)."""
    print("[parity] precision probe: small gate beside the largest load", flush=True)
    lib = CellLibrary()
    d = probe_design(lib)
    tr = probe_trace()
    ref = oracle.run(lib, d, tr, threads=1)
    gp = ref["gate_power"]
    # the probe is what it claims: the inv's power (more events, each below
    # 2e-4 of the largest term) is over 1000x below the buf's switching power
    assert gp[2, 0] > 1e3 * gp[0, 1] > 0
    res, ref = run_both(d, tr, tile_cycles=8)
    compare(res, ref, "precision probe")


def test_small_workload_full_size():
    """The small workload at its full size (2000 gates x 10000 cycles, 10 ns
    clock, 10% DFF, toggle p = 0.2) against the oracle."""
    print("[parity] small_2k full size (2000 gates x 10000 cycles)", flush=True)
    d, tr = S.workload("small_2k")
    assert d.n_gate == 2000 and tr.n_cycle == 10000
    res, ref = run_both(d, tr)
    compare(res, ref, "small_2k")
    assert res.gate_glitch.sum() > 0


@pytest.mark.parametrize("name,gates,cycles,autotune", [("exact_200k", 20_000, 3_000, 1),
                                                           ("skew_1m", 50_000, 2_000, 1)])
def test_scaled_workloads(name, gates, cycles, autotune):
    """The exact-at-scale and skewed-activity workloads with their generators,
    mixes and activity models, at 1/10 and 1/20 of their gate counts (the
    oracle must finish in seconds), with the schedule the autotuner picks on
    them, against the oracle; then in intervals (gtp_run_sliced) bit-identical
    to the resident run."""
    print(f"[parity] {name} scaled to {gates} gates x {cycles} cycles", flush=True)
    d, tr = S.workload(name, scale_gates=gates, scale_cycles=cycles)
    with HipPower(autotune=autotune) as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        res = eng.power()
        sl = eng.power_sliced(tr, max_events=tr.n_event // 5 + 4 * d.n_net)
    ref = oracle.run(d.library, d, tr, threads=0)
    compare(res, ref, name)
    assert res.gate_glitch.sum() > 0
    assert sl.schedule["n_interval"] >= 2
    assert np.array_equal(sl.cycle_power, res.cycle_power)
    assert np.array_equal(sl.gate_power, res.gate_power)


def _float6(v):
    """A float32 value printed as an ostream prints it (6 significant digits)."""
    return float("%g" % float(np.float32(v)))


def test_writers_from_a_gpu_run(tmp_path):
    """cuda_cell_power.txt and the per-cycle waveform file written from a GPU
    run: every printed value equals the oracle's value printed the same way
    (float, 6 significant digits; power/cuda/CudaPower.cu:3134-3165,
    include/sta/PowerClass.hh:123-149), up to the parity bar and one unit in
    the 6th digit."""
    print("[parity] writers from a GPU run", flush=True)
    d = S.make_design(300, dff_fraction=0.1, clock_ns=2.0, seed=61)
    tr = d.trace(200, toggle_p=0.25, seed=62)
    names = [f"top/u{g}" for g in range(d.n_gate)]
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        res = eng.power()
        eng.print_cell_res(tmp_path / "cuda_cell_power.txt", names)
        eng.print_per_cycle_results(tmp_path / "cuda_per_cycle_waveform.txt", "top")
    ref = oracle.run(d.library, d, tr, threads=0)
    G, Cn = d.n_gate, res.n_cycle

    def close(printed, exact, scale):
        p6 = _float6(exact)
        ulp6 = 10.0 ** (np.floor(np.log10(abs(p6))) - 5) if p6 != 0 else 0.0
        return abs(printed - p6) <= 1e-6 * scale + ulp6 + 1e-30

    lines = (tmp_path / "cuda_cell_power.txt").read_text().splitlines()
    assert lines[0].split("\t") == ["cell", "Internal_Power", "Glitch_Internal_Power", "Switching_Power",
                                    "Glitch_Switching_Power", "Leakage_Power", "Total_Power", "Standard_Cell"]
    assert len(lines) == G + 1
    for g, line in enumerate(lines[1:]):
        f = line.split("\t")
        assert f[0] == names[g] and f[7] == d.library.full_names[d.gate_cell[g]]
        for k in range(5):
            assert close(float(f[1 + k]), ref["gate_power"][k, g], ref["gate_abs"][k, g]), (g, k, f)
        tot = ref["gate_power"][:, g].sum()
        assert close(float(f[6]), tot, ref["gate_abs"][:, g].sum()), (g, f)
    t = (tmp_path / "cuda_per_cycle_waveform.txt").read_text().splitlines()
    titles = ["Per cycle leakage power: ", "Per cycle internal power: ", "Per cycle glitch internal power: ",
              "Per cycle switching power: ", "Per cycle glitch switching power: "]
    assert [t[2 * w] for w in range(5)] == titles and t[10] == "End of per cycle power" and len(t) == 11
    for w in range(5):
        f = t[2 * w + 1].split("\t")
        assert f[0] == "top" and f[-1] == "" and len(f) == Cn + 2
        for c in range(Cn):
            assert close(float(f[1 + c]), ref["cycle_power"][w, c], ref["cycle_abs"][w, c]), (w, c, f[1 + c])
