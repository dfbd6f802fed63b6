"""gtp_mlp_predict_noalloc (the allocation-free inference) against
gtp_mlp_predict through the C ABI (GPU): two calls in a row, the second with
fewer rows, on a trained-looking model (random parameters and output
standardisation).  Both run the same kernels on the same rows, so the bar is
bit-equality; the first call spans several chunks of the 16 MiB scratch."""
import ctypes as C

import numpy as np
import pytest

from gtpower_amd import abi
from gtpower_amd.surrogate import MlpSurrogate

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    f = lib.gtp_mlp_predict_noalloc
    f.argtypes = [C.c_void_p, abi.p_u32, C.c_int64, C.POINTER(C.c_float)]
    f.restype = C.c_int32
    return f


def test_noalloc_predict_equals_predict():
    print("[no-alloc] gtp_mlp_predict_noalloc vs gtp_mlp_predict: 25000 rows, then 700", flush=True)
    mlp = MlpSurrogate(n_in=4096, n_hidden=256, seed=5)
    rng = np.random.default_rng(11)
    p, _ = mlp.get_params()
    mlp.set_params(p * 1.5 + rng.normal(0, 1e-3, p.size).astype(np.float32), np.array([0.37, 0.002], np.float32))
    f = _noalloc(mlp._lib)
    for n in (25000, 700):
        bits = rng.integers(0, 2 ** 32, size=(n, 4096 // 32), dtype=np.uint32)
        want = mlp.predict(bits)
        got = np.full(n, np.nan, np.float32)
        st = f(mlp._h, abi.ptr(bits, C.c_uint32), n, got.ctypes.data_as(C.POINTER(C.c_float)))
        assert st == 0, mlp._lib.gtp_mlp_last_error(mlp._h).decode()
        assert np.array_equal(got, want), f"n={n}: max |diff| {np.max(np.abs(got - want))}"
    mlp.close()


def test_noalloc_rejects_a_model_too_large_for_the_scratch():
    # 2^20 inputs: one row's feature words alone are 128 KiB, so 128 rows
    # exceed the 16 MiB scratch
    print("[no-alloc] oversized model refused", flush=True)
    mlp = MlpSurrogate(n_in=1 << 20, n_hidden=128, seed=1)
    f = _noalloc(mlp._lib)
    bits = np.zeros((4, (1 << 20) // 32), np.uint32)
    out = np.zeros(4, np.float32)
    st = f(mlp._h, abi.ptr(bits, C.c_uint32), 4, out.ctypes.data_as(C.POINTER(C.c_float)))
    assert st == -6 and b"scratch" in mlp._lib.gtp_mlp_last_error(mlp._h)
    mlp.close()
