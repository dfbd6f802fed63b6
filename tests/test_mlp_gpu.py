"""Surrogate kernels on the GPU, against the FP64 oracle (oracle/mlp.py) and
brute force.

Tolerances (written here):
  * GEMM layouts: exact (inputs are small multiples of 1/4, so every FP32
    partial sum is exact);
  * forward: |gpu - oracle| <= 1e-4 * std(oracle) + 1e-6 per row, the oracle
    applying the same BF16 rounding points in FP64;
  * gradients: per parameter block, ||gpu - oracle|| <= 2e-3 * ||oracle||;
    loss within 1e-4 relative;
  * toggle bits: bit-exact vs a brute-force scan of the trace; correlation
    within 1e-5 of numpy's Pearson correlation of those bits;
  * judge tiles: bit-exact equality with the full run on the same cycles.
"""
import numpy as np
import pytest

from oracle import mlp as OM
from gtpower_amd import abi
from gtpower_amd import synthetic as S
from gtpower_amd.power import HipPower
from gtpower_amd.surrogate import MlpSurrogate, SurrogatePower, cycle_total, nrmse

pytestmark = pytest.mark.gpu


def _bf16_bits(x):
    f = np.ascontiguousarray(x, dtype=np.float32)
    return (f.view(np.uint32) >> 16).astype(np.uint16)


def _pack_rows(b):
    """[R][n] 0/1 -> [R][n/32] uint32, bit k of a row = bit k%32 of word k/32."""
    R, n = b.shape
    w = b.reshape(R, n // 32, 32).astype(np.uint64) << np.arange(32, dtype=np.uint64)
    return w.sum(axis=2).astype(np.uint32)


def _gemm(a_mode, b_mode, A, B, M, N, K):
    import ctypes as C
    lib = abi.load_engine()
    out = np.zeros((M, N), np.float32)
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    st = lib.gtp_mlp_gemm_test(a_mode, b_mode, M, N, K, A.ctypes.data, A.nbytes, B.ctypes.data, B.nbytes,
                               abi.ptr(out, C.c_float))
    assert st == 0, st
    return out


@pytest.mark.parametrize("a_mode,b_mode", [(0, 0), (0, 1), (1, 0), (1, 2), (2, 1)])
def test_gemm_layouts_exact(a_mode, b_mode):
    print(f"[mlp] gemm a_mode={a_mode} b_mode={b_mode}", flush=True)
    rng = np.random.default_rng(10 * a_mode + b_mode)
    M, N, K = (208 if a_mode == 1 else 200), 256, 96
    if a_mode == 2:
        Ab = (rng.random((M, K)) < 0.4).astype(np.float64)
        A_arg = _pack_rows(Ab)
    else:
        Ab = rng.integers(-8, 9, (M, K)) / 4.0
        A_arg = _bf16_bits(Ab if a_mode == 0 else Ab.T)
    if b_mode == 2:
        Bb = (rng.random((K, N)) < 0.4).astype(np.float64)
        B_arg = _pack_rows(Bb)
    else:
        Bb = rng.integers(-8, 9, (K, N)) / 4.0
        B_arg = _bf16_bits(Bb if b_mode == 0 else Bb.T)
    got = _gemm(a_mode, b_mode, A_arg, B_arg, M, N, K)
    want = Ab @ Bb
    assert np.array_equal(got.astype(np.float64), want), np.argwhere(got != want)[:5]


def _data(n_rows, n_in, seed, p=0.3):
    rng = np.random.default_rng(seed)
    Xb = (rng.random((n_rows, n_in)) < p).astype(np.float64)
    return Xb, _pack_rows(Xb)


def test_forward_matches_oracle():
    n_in, H, n = 256, 128, 300
    print(f"[mlp] forward n_in={n_in} H={H} rows={n}", flush=True)
    m = MlpSurrogate(n_in, H, seed=4)
    p, _ = m.get_params()
    m.set_params(p, np.array([0.0, 1.0], np.float32))
    Xb, bits = _data(n, n_in, 1)
    got = m.predict(bits).astype(np.float64)
    want, _ = OM.forward(p.astype(np.float64), Xb, n_in, H, bf16_points=True)
    tol = 1e-4 * want.std() + 1e-6
    assert np.all(np.abs(got - want) <= tol), np.abs(got - want).max() / want.std()
    m.close()


def test_gradients_match_oracle():
    n_in, H, n = 256, 128, 256
    print(f"[mlp] gradients n_in={n_in} H={H} rows={n}", flush=True)
    m = MlpSurrogate(n_in, H, seed=6)
    p, _ = m.get_params()
    Xb, bits = _data(n, n_in, 2)
    rng = np.random.default_rng(3)
    y = rng.standard_normal(n).astype(np.float32)
    # without dropout, then with the training-time input dropout of one seed
    for seed in (0, 0x1234567):
        loss, g = m.gradients(bits, y, drop_seed=seed)
        X = Xb if seed == 0 else OM.dropped_inputs(bits, n_in, seed)
        lref, gref = OM.loss_and_grad(p.astype(np.float64), X, y.astype(np.float64), n_in, H, bf16_points=True)
        assert abs(loss - lref) <= 1e-4 * lref
        sizes = [H * n_in, H, H * H, H, H, 1]
        names = ["W1", "b1", "W2", "b2", "w3", "b3"]
        o = 0
        for nm, sz in zip(names, sizes):
            a, b = g[o:o + sz].astype(np.float64), gref[o:o + sz]
            err = np.linalg.norm(a - b)
            assert err <= 2e-3 * np.linalg.norm(b) + 1e-12, f"seed {seed} {nm}: rel err {err / np.linalg.norm(b):.3g}"
            o += sz
    m.close()


def test_training_fits_a_linear_target():
    n_in, H, n = 512, 128, 8192
    print(f"[mlp] train n_in={n_in} H={H} rows={n}", flush=True)
    rng = np.random.default_rng(11)
    Xb, bits = _data(n, n_in, 5, p=0.2)
    w = rng.standard_normal(n_in) * (rng.random(n_in) < 0.3)
    y = (Xb @ w + 50.0).astype(np.float32)
    m = MlpSurrogate(n_in, H, seed=2)
    losses = m.train(bits, y, epochs=12, batch=256, lr=2e-3, seed=1)
    assert losses[-1, 0] < 0.1 * losses[0, 0]
    assert np.isnan(losses[:, 1]).all()  # no validation rows
    Xt, bt = _data(2048, n_in, 6, p=0.2)
    pred = m.predict(bt)
    yt = Xt @ w + 50.0
    assert np.sqrt(np.mean((pred - yt) ** 2)) < 0.2 * yt.std()  # held-out rows
    m.close()


@pytest.fixture(scope="module")
def engine_small():
    d = S.make_design(3000, dff_fraction=0.15, clock_ns=2.0, seed=21)
    tr = d.trace(3000, toggle_p=0.15, seed=22)
    eng = HipPower()
    eng.set_design(d.library, d)
    eng.read_power_activities(tr)
    yield d, tr, eng
    eng.close()


def test_toggle_features_brute_force(engine_small):
    d, tr, eng = engine_small
    print("[mlp] toggle features vs brute force", flush=True)
    rng = np.random.default_rng(8)
    N = d.n_net
    feats = np.sort(rng.choice(2 * N, 256, replace=False)).astype(np.int32)  # both kinds
    assert (feats < N).any() and (feats >= N).any()
    c0, n = 100, 2500
    bits = eng.toggle_features(feats, c0, n)
    T = tr.ticks_per_cycle
    want = np.zeros((n, 256), np.uint8)
    off = tr.net_event_offset
    for k, f in enumerate(feats):
        net, rise = f % N, f >= N
        sl = slice(off[net] + 1, off[net + 1])  # index >= 1: value changes
        t, v = tr.event_time[sl], tr.event_value[sl]
        if rise:
            t = t[v == 1]
        c = t // T - c0
        c = c[(c >= 0) & (c < n)]
        want[c, k] = 1
    assert np.array_equal(bits, _pack_rows(want))


@pytest.mark.parametrize("c0,n", [(0, None), (100, 2500), (2000, 900)])
def test_net_weight_sum_brute_force(engine_small, c0, n):
    """The surrogate's base term: per cycle, sum of weight x value changes of
    the selected nets (windows of 2048 cycles, sub-ranges crossing them),
    exact to the weights' 2^-20 quantum."""
    d, tr, eng = engine_small
    print(f"[mlp] net weight sums vs brute force, cycles {c0} +{n}", flush=True)
    rng = np.random.default_rng(9)
    N = d.n_net
    nets = np.sort(rng.choice(N, min(N, 700), replace=False)).astype(np.int32)
    w = rng.uniform(-1e-12, 3e-12, nets.size)
    w[::50] = 0.0
    got, _ = eng.net_weight_sum(nets, w, c0, n)
    n = tr.n_cycle - c0 if n is None else n
    kq = int(np.floor(20.0 - np.log2(np.abs(w).max() * 1.000001)))
    wq = np.rint(np.ldexp(w, kq)).astype(np.int64)
    T = tr.ticks_per_cycle
    off = tr.net_event_offset
    acc = np.zeros(n, np.int64)
    for k, net in enumerate(nets):
        c = tr.event_time[off[net] + 1:off[net + 1]] // T - c0   # index >= 1: value changes
        c = c[(c >= 0) & (c < n)]
        np.add.at(acc, c, wq[k])
    assert np.array_equal(got, np.ldexp(acc.astype(np.float64), -kq))
    assert acc.any()


def test_correlation_matches_numpy(engine_small):
    d, tr, eng = engine_small
    print("[mlp] correlation vs numpy", flush=True)
    res = eng.power(want=("cycle",))
    y = cycle_total(res.cycle_power)
    mask = np.zeros(y.size, np.uint8)
    mask[::3] = 1
    corr = eng.feature_correlation(y, mask)
    N = d.n_net
    assert corr.shape == (2 * N,)
    nets = np.arange(0, 1024, dtype=np.int32)
    for feats in (nets, nets + N):  # value changes, rises
        X = OM.unpack_bits(eng.toggle_features(feats), len(feats))[mask.astype(bool)]
        ym = y[mask.astype(bool)]
        xs = X.std(axis=0)
        want = np.where(xs > 0, ((X - X.mean(0)) * (ym - ym.mean())[:, None]).mean(0) / (xs * ym.std() + 1e-300),
                        0.0)
        assert np.allclose(corr[feats], want, rtol=0, atol=1e-5)


def test_judge_tiles_equal_full_run(engine_small):
    d, tr, eng = engine_small
    print("[mlp] run_tiles vs run", flush=True)
    full = eng.power(want=("cycle",))
    tc = full.schedule["tile_cycles"]
    n_tile = full.schedule["n_tile"]
    tiles = np.array(sorted({0, 3, n_tile - 1, n_tile // 2}), np.int32)
    cyc, waves, ms = eng.power_tiles(tc, tiles)
    assert np.array_equal(waves, full.cycle_power[:, cyc])
    # another tile length (the full run's values do not depend on it)
    cyc2, waves2, _ = eng.power_tiles(64, np.array([1, 5], np.int32))
    assert np.allclose(waves2, full.cycle_power[:, cyc2], rtol=1e-9, atol=1e-30)
    with pytest.raises(abi.GtpError):
        eng.power_tiles(tc, np.array([n_tile], np.int32))


def test_judge_tiles_before_any_full_run():
    """Tiles evaluated on a fresh trace (only their own events and the ones
    next to them annotated), adjacent and edge tiles included, then the same
    engine after a different trace: bit-identical to the full run each time."""
    d = S.make_design(2000, dff_fraction=0.15, clock_ns=2.0, seed=31)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        for seed in (32, 33):
            print(f"[mlp] fresh-trace run_tiles seed {seed}", flush=True)
            tr = d.trace(2500, toggle_p=0.2, seed=seed)
            eng.read_power_activities(tr)
            tc = 64
            n_tile = -(-tr.n_cycle // tc)
            tiles = np.array(sorted({0, 3, 4, n_tile // 2, n_tile - 1}), np.int32)
            cyc, waves, _ = eng.power_tiles(tc, tiles)
            full = eng.power(want=("cycle",))
            assert np.allclose(waves, full.cycle_power[:, cyc], rtol=1e-9, atol=1e-30)
            if full.schedule["tile_cycles"] == tc:
                assert np.array_equal(waves, full.cycle_power[:, cyc])


def test_surrogate_gain_and_fallback():
    """What the trained surrogate gains: per-cycle power of a fresh trace at a
    fraction of the exact engine's device time, within 2% NRMSE (RMSE / mean
    power) and at most 0.8x the NRMSE of a constant predictor at the true mean;
    the judge falls back to the exact answer when its threshold is not met, and
    when a model predicts only the mean, however loose the absolute bar."""
    print("[mlp] surrogate end to end: 50000 gates, 20000 + 6000 cycles", flush=True)
    d = S.make_design(50000, dff_fraction=0.15, clock_ns=2.0, seed=31)
    train = d.trace(20000, toggle_p=0.15, seed=32)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(train)
        sp = SurrogatePower(MlpSurrogate(4096, 256, seed=3), threshold=0.02, judge_fraction=0.05)
        sp.fit(eng)
        del train
        fresh = d.trace(6000, toggle_p=0.15, seed=33)
        eng.read_power_activities(fresh)
        exact = eng.power(want=("cycle",))
        y = cycle_total(exact.cycle_power)
        out = sp.evaluate(eng)
        full_err = nrmse(out.cycle_total, y)
        r2 = 1.0 - np.mean((out.cycle_total - y) ** 2) / y.var()
        print(f"[mlp] judged nrmse={out.nrmse:.4f} full nrmse={full_err:.4f} r2={r2:.3f} "
              f"constant-predictor nrmse={y.std() / y.mean():.4f} source={out.source} "
              f"surrogate_ms={out.surrogate_ms:.3f} judge_ms={out.judge_ms:.3f} exact_ms={exact.kernel_ms:.3f}",
              flush=True)
        const = y.std() / y.mean()
        assert out.source == "surrogate"
        assert full_err <= 0.02
        assert full_err <= 0.8 * const
        assert out.nrmse <= sp.rel_gain * out.spread
        assert out.surrogate_ms + out.judge_ms < exact.kernel_ms
        # the first cycles (tile 0) come back exactly
        assert np.array_equal(out.cycle_total[:sp.judge_tile], y[:sp.judge_tile])
        # a threshold the surrogate cannot meet: the exact answer comes back
        sp.threshold = 1e-9
        fb = sp.evaluate(eng)
        assert fb.source == "exact"
        assert np.array_equal(fb.cycle_total, y)
        # a model that predicts the trace's mean power: within an absolute
        # bar of 3%, but no better than the judged cycles' own spread -> exact
        sp.threshold = 0.03
        sp.base_nets = None
        p, _ = sp.mlp.get_params()
        sp.mlp.set_params(np.zeros_like(p), np.array([y.mean(), 1.0], np.float32))
        mean_model = sp.evaluate(eng)
        print(f"[mlp] mean model: judged nrmse={mean_model.nrmse:.4f} spread={mean_model.spread:.4f} "
              f"source={mean_model.source}", flush=True)
        assert mean_model.nrmse <= 0.03
        assert mean_model.source == "exact"
        assert np.array_equal(mean_model.cycle_total, y)
        sp.mlp.close()
