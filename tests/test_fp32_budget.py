"""FP32-versus-FP64 error budget of the CPU restatement (CPU).

The reference's CPU path computes energies, per-gate sums and per-thread
per-cycle waveforms in float (util/Types.hh:3-18); the oracle computes them in
double.  The oracle's FP32 build (liboracle_f32.so, the same code with
ORACLE_REAL=float) shows how far the reference's own arithmetic moves from the
FP64 values: counts are integers and must agree exactly, and the values stay
within an FP32 budget, measured here against the same term-magnitude scale
as the engine's 1e-6 parity bar (tests/test_engine_gpu.py)."""
import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S


def rel_err(a, b, scale):
    return np.abs(a - b) / np.maximum(scale, 1e-300)


@pytest.mark.parametrize("n_gate,n_cycle,seed", [(300, 400, 3), (1500, 1000, 4)])
def test_fp32_build_within_its_budget(n_gate, n_cycle, seed):
    d = S.make_design(n_gate, dff_fraction=0.1, clock_ns=10.0, seed=seed)
    tr = d.trace(n_cycle, toggle_p=0.2, seed=seed + 100)
    ref = oracle.run(d.library, d, tr, threads=4)
    f32 = oracle.run(d.library, d, tr, threads=4, fp32=True)
    # integer counts do not depend on the value type
    assert np.array_equal(ref["gate_toggle_x2"], f32["gate_toggle_x2"])
    assert np.array_equal(ref["gate_glitch"], f32["gate_glitch"])
    ge = rel_err(f32["gate_power"], ref["gate_power"], ref["gate_abs"])
    ce = rel_err(f32["cycle_power"], ref["cycle_power"], ref["cycle_abs"])
    print(f"[fp32 budget] {n_gate} gates x {n_cycle} cycles: per-gate max {ge.max():.3e}, "
          f"per-cycle max {ce.max():.3e} (of the term-magnitude scale)")
    # FP32 arithmetic: a few ulps of float per term, growing with the number
    # of float additions into one gate or one cycle
    assert ge.max() < 1e-4
    assert ce.max() < 1e-4
    # the FP32 build really computes in FP32 (it is not the FP64 oracle again)
    assert ge.max() > 1e-9 or ce.max() > 1e-9


def test_fp32_build_matches_hand_case():
    """A constant-state gate leaks exactly its table value in every cycle in
    both builds (no rounding in a single FP32 value times whole cycles)."""
    d = S.make_design(50, dff_fraction=0.0, clock_ns=10.0, seed=9)
    tr = d.trace(20, toggle_p=0.0, seed=10)
    ref = oracle.run(d.library, d, tr, threads=1)
    f32 = oracle.run(d.library, d, tr, threads=1, fp32=True)
    assert np.allclose(f32["gate_power"][4], ref["gate_power"][4], rtol=1e-6, atol=0)
