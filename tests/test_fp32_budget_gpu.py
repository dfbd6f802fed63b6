"""The engine against the reference's own FP32 arithmetic (GPU).

The engine evaluates every term in FP32 and sums exact integer quanta; the
reference sums FP32 values (util/Types.hh:3-18), which the oracle's FP32 build
reproduces.  Measured against the FP64 oracle on the same term-magnitude
scale, the engine's per-gate error stays within the 1e-6 parity bar and below
the FP32 build's own error."""
import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.power import HipPower

pytestmark = pytest.mark.gpu


def test_engine_error_below_the_fp32_reference_error():
    print("[fp32 budget] 1500 gates x 1000 cycles: engine vs FP32 build vs FP64 oracle", flush=True)
    d = S.make_design(1500, dff_fraction=0.1, clock_ns=10.0, seed=4)
    tr = d.trace(1000, toggle_p=0.2, seed=104)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        res = eng.power()
    ref = oracle.run(d.library, d, tr, threads=0)
    f32 = oracle.run(d.library, d, tr, threads=0, fp32=True)
    scale_g = np.maximum(ref["gate_abs"], 1e-300)
    scale_c = np.maximum(ref["cycle_abs"], 1e-300)
    eng_g = (np.abs(res.gate_power - ref["gate_power"]) / scale_g).max()
    f32_g = (np.abs(f32["gate_power"] - ref["gate_power"]) / scale_g).max()
    eng_c = (np.abs(res.cycle_power - ref["cycle_power"]) / scale_c).max()
    f32_c = (np.abs(f32["cycle_power"] - ref["cycle_power"]) / scale_c).max()
    print(f"[fp32 budget] per gate: engine {eng_g:.3e}, FP32 build {f32_g:.3e}; "
          f"per cycle: engine {eng_c:.3e}, FP32 build {f32_c:.3e}", flush=True)
    assert eng_g <= 1e-6 and eng_c <= 1e-6
    assert eng_g < f32_g
