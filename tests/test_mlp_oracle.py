"""Pinning of the FP64 surrogate oracle (CPU).

The oracle MLP (oracle/mlp.py) is checked against what fixes it independently
of its own formulas:
  * bf16 rounding: hand-worked ties and torch's float32 -> bfloat16 conversion;
  * GELU: closed-form values and torch's erf GELU, its derivative by central
    differences;
  * gradients: central finite differences of the FP64 loss on every parameter,
    and torch autograd of the same network in float64;
  * bit unpacking: hand-built words.
"""
import numpy as np
import pytest
import torch

from oracle import mlp as M


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 3 * 2 ** -8), 1 + 2 ** -7, 0.0, 3.0e38, 1e-40])
    want = np.array([1.0, 1.0, 1 + 2 ** -6, -(1 + 2 ** -6), 1 + 2 ** -7, 0.0, None, None], dtype=object)
    got = M.bf16(x)
    for g, w in zip(got[:6], want[:6]):
        assert g == w
    # agrees with torch's conversion everywhere on a spread of magnitudes
    rng = np.random.default_rng(0)
    v = (rng.standard_normal(100000) * np.exp(rng.uniform(-20, 20, 100000))).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(v), ref)


def test_gelu_closed_forms_and_torch():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu(np.array([10.0]))[0] - 10.0) < 1e-12
    assert abs(M.gelu(np.array([-10.0]))[0]) < 1e-12
    assert abs(M.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15  # Phi(1)
    z = np.linspace(-6, 6, 1001)
    ref = torch.nn.functional.gelu(torch.from_numpy(z)).numpy()
    assert np.allclose(M.gelu(z), ref, rtol=0, atol=1e-14)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(z), fd, rtol=0, atol=1e-8)


def test_unpack_bits():
    w = np.array([[0x00000001, 0x80000000], [0xFFFFFFFF, 0x00000002]], np.uint32)
    X = M.unpack_bits(w, 64)
    assert X.shape == (2, 64)
    assert X[0].nonzero()[0].tolist() == [0, 63]
    assert X[1, :32].sum() == 32 and X[1, 32:].nonzero()[0].tolist() == [1]


def _tiny(seed=3, n_in=64, H=16, M_=24):
    rng = np.random.default_rng(seed)
    n_param = H * n_in + H + H * H + H + H + 1
    p = rng.standard_normal(n_param) * 0.3
    X = (rng.random((M_, n_in)) < 0.3).astype(np.float64)
    y = rng.standard_normal(M_)
    return p, X, y, n_in, H


def test_gradient_matches_finite_differences():
    p, X, y, n_in, H = _tiny()
    _, g = M.loss_and_grad(p, X, y, n_in, H, bf16_points=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, X, y, n_in, H, bf16_points=False)
        lm, _ = M.loss_and_grad(p - e, X, y, n_in, H, bf16_points=False)
        fd[i] = (lp - lm) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-8), np.abs(g - fd).max()


def test_gradient_matches_torch_autograd():
    p, X, y, n_in, H = _tiny(seed=5)
    loss, g = M.loss_and_grad(p, X, y, n_in, H, bf16_points=False)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    o = 0
    W1 = pt[o:o + H * n_in].view(H, n_in); o += H * n_in
    b1 = pt[o:o + H]; o += H
    W2 = pt[o:o + H * H].view(H, H); o += H * H
    b2 = pt[o:o + H]; o += H
    w3 = pt[o:o + H]; o += H
    b3 = pt[o]
    Xt, yt = torch.from_numpy(X), torch.from_numpy(y)
    h1 = torch.nn.functional.gelu(Xt @ W1.T + b1)
    h2 = torch.nn.functional.gelu(h1 @ W2.T + b2)
    lt = ((h2 @ w3 + b3 - yt) ** 2).mean()
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12 * max(1.0, loss)
    assert np.allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_bf16_points_change_little():
    p, X, y, n_in, H = _tiny(seed=9)
    l0, g0 = M.loss_and_grad(p, X, y, n_in, H, bf16_points=False)
    l1, g1 = M.loss_and_grad(p, X, y, n_in, H, bf16_points=True)
    assert abs(l1 - l0) < 0.05 * l0
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)
    assert not np.array_equal(g0, g1)  # the rounding points are applied


def test_select_features_ranking():
    from gtpower_amd.surrogate import select_features
    corr = np.array([0.1, -0.9, 0.5, 0.0, 0.9, np.nan, -0.5, 0.2], np.float32)
    # |corr|: 0.9 (1), 0.9 (4), 0.5 (2), 0.5 (6) ...; ties by lower index
    assert select_features(corr, 3).tolist() == [1, 2, 4]
    assert select_features(corr, 4).tolist() == [1, 2, 4, 6]


def test_nrmse_definition():
    from gtpower_amd.surrogate import nrmse
    exact = np.array([1.0, 2.0, 3.0, 2.0])
    pred = exact + np.array([0.1, -0.1, 0.1, -0.1])
    assert abs(nrmse(pred, exact) - 0.1 / 2.0) < 1e-15


def test_keep_mask_known_values_and_balance():
    # the finaliser on hand-computed inputs: seed=1,row=0,word=0 -> h = 1
    h = 1
    h ^= h >> 16
    h = (h * 0x7FEB352D) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x846CA68B) & 0xFFFFFFFF
    h ^= h >> 16
    assert int(M.keep_mask(1, 0, 0)) == h
    # about half of the bits are kept, independently of row and word
    km = M.keep_mask(12345, np.arange(2000)[:, None], np.arange(64)[None, :])
    frac = np.unpackbits(km.view(np.uint8)).mean()
    assert abs(frac - 0.5) < 0.005


def test_keep_mask_matches_engine_export():
    from gtpower_amd import abi
    try:
        lib = abi.load_engine()
    except (RuntimeError, OSError):
        pytest.skip("engine library not built")
    for seed, r, w in [(1, 0, 0), (0xDEADBEEF, 17, 3), (7, 123456, 127), (0xFFFFFFFF, 0xFFFFFFFF, 5)]:
        assert lib.gtp_keep_mask(seed, r, w) == int(M.keep_mask(seed, r, w))
