"""Learned per-cycle power surrogate, judged by the exact engine.

No counterpart in the reference: a new method on top of its time-based path.
The exact engine (``HipPower.power``) produces per-cycle power; a small MLP
learns it from per-cycle toggle bits -- (net, kind) features, kind 0 a value
change of the net in the cycle, kind 1 a change to 1 -- of the features whose
bits correlate most with it; the exact engine then recomputes a sample of the cycles
(``HipPower.power_tiles``) and the surrogate's answer is replaced by the exact
one when the sampled error is too large.

    mlp = MlpSurrogate(n_in=4096, n_hidden=256)
    sp = SurrogatePower(mlp)
    sp.fit(engine, train_mask)            # engine holds a loaded trace
    engine.read_power_activities(fresh)   # a new trace of the same design
    out = sp.evaluate(engine)             # surrogate + judge (+ fallback)

Everything numeric runs on the device through libgtpower_hip.so: features
(``gtp_toggle_features`` / ``gtp_mlp_predict_trace``), correlation
(``gtp_feature_correlation``), training (``gtp_mlp_train``: BF16 matrix-core
forward and backward kernels, FP32 Adam) and the judge (``gtp_run_tiles``).

Definitions used by the judge:
  target  y_c   = leakage + internal + switching power of cycle c (W; internal
                  and switching include their glitch parts);
  NRMSE         = sqrt(mean_c (y^_c - y_c)^2) / mean_c(y_c) over the judged cycles.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


def cycle_total(cycle_power):
    """Per-cycle total power from the 5 waveforms (leakage, internal, glitch
    internal, switching, glitch switching): leakage + internal + switching."""
    cp = np.asarray(cycle_power)
    return cp[0] + cp[1] + cp[3]


def nrmse(pred, exact):
    exact = np.asarray(exact, dtype=np.float64)
    d = np.asarray(pred, dtype=np.float64) - exact
    return float(np.sqrt(np.mean(d * d)) / np.mean(exact))


class MlpSurrogate:
    """The device MLP (n_in toggle bits -> n_hidden -> n_hidden -> 1, GELU)."""

    def __init__(self, n_in=4096, n_hidden=256, seed=1, device=0):
        self._lib = abi.load_engine()
        cfg = abi.GtpMlpConfig(device, n_in, n_hidden, 0, seed)
        h = C.c_void_p()
        st = self._lib.gtp_mlp_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise abi.GtpError(st, "gtp_mlp_create", f"n_in={n_in} n_hidden={n_hidden}")
        self._h = h
        self.n_in, self.n_hidden = n_in, n_hidden
        self.n_param = int(self._lib.gtp_mlp_param_count(h))

    def _check(self, st, where):
        if st != 0:
            raise abi.GtpError(st, where, self._lib.gtp_mlp_last_error(self._h).decode())

    def get_params(self):
        p = np.zeros(self.n_param, np.float32)
        ms = np.zeros(2, np.float32)
        self._check(self._lib.gtp_mlp_get_params(self._h, abi.ptr(p, C.c_float), abi.ptr(ms, C.c_float)),
                    "gtp_mlp_get_params")
        return p, ms

    def set_params(self, params, y_mean_std=None):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.n_param
        ms = None if y_mean_std is None else np.ascontiguousarray(y_mean_std, dtype=np.float32)
        self._check(self._lib.gtp_mlp_set_params(self._h, abi.ptr(p, C.c_float), abi.ptr(ms, C.c_float)),
                    "gtp_mlp_set_params")

    def gradients(self, bits, y_std, drop_seed=0):
        """(loss, grads) of the mean squared error on standardised targets
        (drop_seed != 0: with the training-time input dropout of that seed)."""
        bits = np.ascontiguousarray(bits, dtype=np.uint32)
        y = np.ascontiguousarray(y_std, dtype=np.float32)
        g = np.zeros(self.n_param, np.float32)
        loss = C.c_float()
        self._check(self._lib.gtp_mlp_gradients(self._h, abi.ptr(bits, C.c_uint32), abi.ptr(y, C.c_float), len(y),
                                                drop_seed, abi.ptr(g, C.c_float), C.byref(loss)),
                    "gtp_mlp_gradients")
        return float(loss.value), g

    def train(self, bits, y, epochs=30, batch=256, lr=5e-4, weight_decay=0.0, val_rows=0, input_dropout=False,
              seed=7):
        """Train on the rows of (bits, y); the last val_rows rows are held out
        and the best epoch on them is kept.  Returns [epochs, 2]: mean training
        loss and validation loss per epoch (standardised targets)."""
        bits = np.ascontiguousarray(bits, dtype=np.uint32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        tc = abi.GtpTrainConfig(batch, epochs, lr, weight_decay, seed, val_rows, 1 if input_dropout else 0)
        el = np.zeros(2 * epochs, np.float32)
        self._check(self._lib.gtp_mlp_train(self._h, abi.ptr(bits, C.c_uint32), abi.ptr(y, C.c_float), len(y),
                                            C.byref(tc), abi.ptr(el, C.c_float)), "gtp_mlp_train")
        return el.reshape(epochs, 2)

    def predict(self, bits):
        bits = np.ascontiguousarray(bits, dtype=np.uint32)
        out = np.zeros(bits.shape[0], np.float32)
        self._check(self._lib.gtp_mlp_predict(self._h, abi.ptr(bits, C.c_uint32), bits.shape[0],
                                              abi.ptr(out, C.c_float)), "gtp_mlp_predict")
        return out

    def predict_trace(self, engine, feats, cycle_begin=0, n_cycle=None):
        """Predictions for the trace loaded in `engine` from the features
        `feats` (ids kind * n_net + net); returns (pred, device ms)."""
        nets = np.ascontiguousarray(feats, dtype=np.int32)
        assert nets.size == self.n_in
        n = engine.n_cycle - cycle_begin if n_cycle is None else n_cycle
        out = np.zeros(n, np.float32)
        ms = C.c_double()
        self._check(self._lib.gtp_mlp_predict_trace(self._h, engine._h, abi.ptr(nets, C.c_int32), int(cycle_begin),
                                                    int(n), abi.ptr(out, C.c_float), C.byref(ms)),
                    "gtp_mlp_predict_trace")
        return out, ms.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gtp_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def select_features(corr, n_in):
    """Ids of the n_in features with the largest |correlation| (ties: lower
    id first), in increasing id order."""
    a = np.abs(np.nan_to_num(np.asarray(corr, dtype=np.float64)))
    order = np.lexsort((np.arange(a.size), -a))
    return np.sort(order[:n_in]).astype(np.int32)


@dataclass
class SurrogateOutput:
    cycle_total: np.ndarray   # per-cycle total power returned to the caller (W)
    source: str               # "surrogate" or "exact" (fallback)
    nrmse: float              # judged NRMSE of the surrogate
    judged_cycles: int
    surrogate_ms: float       # device time: features + inference
    judge_ms: float           # device time: exact evaluation of the judged tiles
    fallback_ms: float        # device time of the exact fallback (0 if none)


class SurrogatePower:
    """Surrogate with an exact judge and an exact fallback."""

    def __init__(self, mlp, threshold=0.02, judge_fraction=0.01, judge_tile=64, seed=0x47545057):
        # short judge tiles: a tile's gates walk its cycles serially, so many
        # short tiles finish sooner than few long ones
        self.mlp = mlp
        self.threshold = threshold
        self.judge_fraction = judge_fraction
        self.judge_tile = judge_tile
        self.seed = seed
        self.feats = None

    def fit(self, engine, train_mask=None, exact=None, epochs=30, batch=256, lr=5e-4, weight_decay=0.0,
            val_fraction=0.1, input_dropout=True, seed=7):
        """Train on the exact per-cycle power of the cycles with train_mask set
        (all cycles when None) of the trace loaded in `engine`; the last
        val_fraction of those cycles are held out to pick the best epoch.
        `exact` may pass a PowerResult of that trace already computed."""
        if exact is None:
            exact = engine.power(want=("cycle",))
        y = cycle_total(exact.cycle_power)
        mask = np.ones(y.size, bool) if train_mask is None else np.asarray(train_mask, bool)
        rows = np.nonzero(mask)[0]
        n_val = int(val_fraction * len(rows))
        fit_mask = np.zeros(y.size, np.uint8)
        fit_mask[rows[:len(rows) - n_val]] = 1  # features are ranked on the training rows only
        corr = engine.feature_correlation(y, fit_mask)
        self.feats = select_features(corr, self.mlp.n_in)
        bits = engine.toggle_features(self.feats)
        self.train_loss = self.mlp.train(bits[rows], y[rows], epochs=epochs, batch=batch, lr=lr,
                                         weight_decay=weight_decay, val_rows=n_val, input_dropout=input_dropout,
                                         seed=seed)
        return self

    def judge_tiles(self, n_cycle):
        """Tiles recomputed exactly: judge_fraction of the tiles (at least
        one), drawn without replacement with the recorded seed."""
        n_tile = -(-n_cycle // self.judge_tile)
        k = max(1, int(round(self.judge_fraction * n_tile)))
        rng = np.random.default_rng(self.seed)
        return np.sort(rng.choice(n_tile, size=min(k, n_tile), replace=False)).astype(np.int32)

    def evaluate(self, engine):
        """Per-cycle total power of the trace loaded in `engine`."""
        assert self.feats is not None, "fit() first"
        pred, s_ms = self.mlp.predict_trace(engine, self.feats)
        cyc, waves, j_ms = engine.power_tiles(self.judge_tile, self.judge_tiles(engine.n_cycle))
        err = nrmse(pred[cyc], cycle_total(waves))
        if err <= self.threshold:
            return SurrogateOutput(pred.astype(np.float64), "surrogate", err, len(cyc), s_ms, j_ms, 0.0)
        res = engine.power(want=("cycle",))
        return SurrogateOutput(cycle_total(res.cycle_power), "exact", err, len(cyc), s_ms, j_ms, res.total_ms)
