"""Learned per-cycle power surrogate, judged by the exact engine.

No counterpart in the reference: a new method on top of its time-based path.
The exact engine (``HipPower.power``) produces per-cycle power.  The surrogate
is a linear base term -- per net, the dynamic power of its driving gates per
value change (from the exact run's per-gate power), times the net's value
changes in the cycle, summed over the nets that carry most of that term's
variance (``gtp_net_weight_sum``) -- plus a small MLP that learns what the base
leaves from per-cycle toggle bits -- (net, kind) features, kind 0 a value
change of the net in the cycle, kind 1 a change to 1 -- chosen residual-greedily
by their correlation with it.  The exact engine then recomputes a sample of the
cycles (``HipPower.power_tiles``) and the surrogate's answer is replaced by the
exact one when the sampled error is too large in absolute terms or relative to
the sampled cycles' own spread.

    mlp = MlpSurrogate(n_in=4096, n_hidden=256)
    sp = SurrogatePower(mlp)
    sp.fit(engine, train_mask)            # engine holds a loaded trace (exact run inside)
    engine.read_power_activities(fresh)   # a new trace of the same design
    out = sp.evaluate(engine)             # surrogate + judge (+ fallback)

Everything numeric over cycles runs on the device through libgtpower_hip.so:
features (``gtp_toggle_features`` / ``gtp_mlp_predict_trace``), the base term
(``gtp_net_weight_sum``), correlation (``gtp_feature_correlation``), training
(``gtp_mlp_train``: BF16 matrix-core forward and backward kernels, FP32 Adam)
and the judge (``gtp_run_tiles``); the host only fits the base term's two
coefficients and the greedy rounds' ridge steps (512 features each).

Definitions used by the judge:
  target  y_c   = leakage + internal + switching power of cycle c (W; internal
                  and switching include their glitch parts);
  NRMSE         = sqrt(mean_c (y^_c - y_c)^2) / mean_c(y_c) over the judged cycles.
"""
import ctypes as C
import threading
import time
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


def select_features(corr, n_in, exclude=None):
    """Ids of the n_in features with the largest |correlation| (ties: lower
    id first; `exclude`: ids never picked), in increasing id order."""
    a = np.abs(np.nan_to_num(np.asarray(corr, dtype=np.float64)))
    if exclude is not None and len(exclude):
        a[np.asarray(exclude)] = -1.0
    order = np.lexsort((np.arange(a.size), -a))
    return np.sort(order[:n_in]).astype(np.int32)


def unpack_rows(bits, rows):
    """Rows of packed feature bits [n][words] uint32 -> [len(rows)][32 words]
    float32 0/1 (bit k = bit k % 32 of word k / 32)."""
    b = np.ascontiguousarray(np.asarray(bits)[rows], dtype=np.uint32)
    return np.unpackbits(b.view(np.uint8), axis=1, bitorder="little").astype(np.float32)


def select_features_greedy(engine, y, fit_mask, n_in, rounds=8, ridge=1e-3):
    """Residual-greedy feature choice: each round ranks the features by their
    correlation with what the features picked so far leave unexplained (a
    ridge fit of the round's picks on the fit rows), so the later rounds pick
    features that add information instead of near-copies of the first ones
    (toggles of a net and of its fanout are strongly correlated).  n_in / rounds
    features per round (a multiple of 32); increasing id order."""
    per = n_in // rounds
    assert per % 32 == 0 and per * rounds == n_in
    rows = np.nonzero(fit_mask)[0]
    r = np.asarray(y, dtype=np.float64).copy()
    chosen = np.zeros(0, np.int32)
    for _ in range(rounds):
        corr = engine.feature_correlation(r, fit_mask)
        pick = select_features(corr, per, exclude=chosen)
        chosen = np.concatenate([chosen, pick])
        X = unpack_rows(engine.toggle_features(pick), rows).astype(np.float64)
        xm = X.mean(axis=0)
        Xc = X - xm
        t = r[rows] - r[rows].mean()
        G = Xc.T @ Xc
        G[np.diag_indices_from(G)] += ridge * (np.trace(G) / per + 1e-12)
        w = np.linalg.solve(G, Xc.T @ t)
        r[rows] -= Xc @ w          # only the fit rows enter the next correlation
    return np.sort(chosen).astype(np.int32)


def toggle_energy_weights(design, library, gate_power, net_events):
    """Per net, the dynamic power of its driving gates per value change of the
    net (W per change over the trace): each gate's internal and switching
    power (glitch parts included) is shared by the value changes of its
    output nets.  The surrogate's linear base term, sum over nets of weight x
    changes in a cycle, is then proportional to a toggle-count estimate of the
    cycle's dynamic power."""
    G = design.n_gate
    P = np.asarray(gate_power, dtype=np.float64)[:4].sum(axis=0)
    cell = np.asarray(design.gate_cell)
    npin = np.asarray(library.cell_n_pin)[cell].astype(np.int64)
    nin = np.asarray(library.cell_n_in)[cell].astype(np.int64)
    off = np.asarray(design.gate_pin_offset, dtype=np.int64)
    gate_of_pin = np.repeat(np.arange(G), npin)
    local = np.arange(int(off[-1])) - off[gate_of_pin]
    out = local >= nin[gate_of_pin]
    nets = np.asarray(design.pin_net)[out].astype(np.int64)
    g = gate_of_pin[out]
    keep = nets >= 0
    nets, g = nets[keep], g[keep]
    ch = np.maximum(np.asarray(net_events, dtype=np.int64)[nets] - 1, 0).astype(np.float64)
    tot = np.bincount(g, weights=ch, minlength=G)
    share = np.where(tot[g] > 0, P[g] / np.maximum(tot[g], 1.0), 0.0)
    w = np.zeros(design.n_net)
    np.add.at(w, nets, share)
    return w


def select_base_nets(w, net_events, cover=0.995, max_nets=1 << 20):
    """Nets of the base term: the fewest whose w^2 x changes (their share of
    the term's variance, for rare independent toggles) reach `cover` of the
    total; increasing id order."""
    ch = np.maximum(np.asarray(net_events, dtype=np.int64) - 1, 0).astype(np.float64)
    score = np.asarray(w) ** 2 * ch
    order = np.argsort(-score, kind="stable")
    cum = np.cumsum(score[order])
    if cum[-1] <= 0:
        return np.zeros(0, np.int32)
    k = min(int(np.searchsorted(cum, cover * cum[-1])) + 1, max_nets, int((score > 0).sum()))
    return np.sort(order[:k]).astype(np.int32)


@dataclass
class SurrogateOutput:
    cycle_total: np.ndarray   # per-cycle total power returned to the caller (W)
    source: str               # "surrogate" or "exact" (fallback)
    nrmse: float              # judged NRMSE of the surrogate
    judged_cycles: int
    surrogate_ms: float       # device time: features + inference
    judge_ms: float           # device time: exact evaluation of the judged tiles
    fallback_ms: float        # device time of the exact fallback (0 if none)
    spread: float = 0.0       # std / mean of the judged exact cycles (a constant predictor's NRMSE)
    wall_ms: float = 0.0      # host wall time of the evaluation (inference and judge overlapped)


class SurrogatePower:
    """Surrogate with an exact judge and an exact fallback.

    The surrogate's answer stands when its NRMSE on the judged cycles is at
    most `threshold` AND at most `rel_gain` times the judged cycles' own
    spread (std / mean: what predicting their mean would score), so that a
    model that learned nothing beyond the mean falls back to the exact answer
    even when the power varies little from cycle to cycle."""

    def __init__(self, mlp, threshold=0.02, judge_fraction=0.005, judge_tile=64, seed=0x47545057, rel_gain=0.9,
                 selection="greedy", overlap=True, base=True, base_cover=0.99, skip_cycles=4):
        # short judge tiles: a tile's gates walk its cycles serially, so many
        # short tiles finish sooner than few long ones
        self.mlp = mlp
        self.threshold = threshold
        self.rel_gain = rel_gain
        self.judge_fraction = judge_fraction
        self.judge_tile = judge_tile
        self.seed = seed
        self.selection = selection
        self.overlap = overlap
        self.base = base
        self.base_cover = base_cover
        self.skip_cycles = skip_cycles
        self.base_nets = None     # nets of the linear base term, their weights, and its fit y ~ a x sum + b
        self.base_w = None
        self.base_ab = (0.0, 0.0)
        self.feats = None

    def fit(self, engine, train_mask=None, exact=None, epochs=30, batch=256, lr=5e-4, weight_decay=0.0,
            val_fraction=0.1, input_dropout=True, seed=7):
        """Train on the exact per-cycle power of the cycles with train_mask set
        (all cycles when None) of the trace loaded in `engine`; the last
        val_fraction of those cycles are held out to pick the best epoch.
        `exact` may pass a PowerResult of that trace already computed."""
        if exact is None or (self.base and getattr(exact, "gate_power", None) is None):
            exact = engine.power(want=("gate", "cycle"))
        y = cycle_total(exact.cycle_power)
        mask = np.ones(y.size, bool) if train_mask is None else np.asarray(train_mask, bool).copy()
        # the first cycles of a trace (initial values settling) are outliers of
        # both the power and the toggle counts: they would steer the fits
        mask[:self.skip_cycles] = False
        rows = np.nonzero(mask)[0]
        n_val = int(val_fraction * len(rows))
        fit_mask = np.zeros(y.size, np.uint8)
        fit_mask[rows[:len(rows) - n_val]] = 1  # features are ranked on the training rows only
        if self.base:
            # the linear base term, fitted on the training rows; the network
            # learns what it leaves
            ev = engine.net_event_counts()
            w = toggle_energy_weights(engine.design, engine.library, exact.gate_power, ev)
            self.base_nets = select_base_nets(w, ev, self.base_cover)
            self.base_w = w[self.base_nets]
            lin, _ = engine.net_weight_sum(self.base_nets, self.base_w)
            fr = fit_mask.astype(bool)
            A = np.stack([lin[fr], np.ones(int(fr.sum()))], axis=1)
            a, b = np.linalg.lstsq(A, y[fr], rcond=None)[0]
            self.base_ab = (float(a), float(b))
            y = y - (a * lin + b)
        if self.selection == "greedy":
            self.feats = select_features_greedy(engine, y, fit_mask, self.mlp.n_in)
        else:
            self.feats = select_features(engine.feature_correlation(y, fit_mask), self.mlp.n_in)
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

    def accepts(self, err, spread):
        return err <= self.threshold and err <= self.rel_gain * spread

    def _base(self, engine):
        if self.base_nets is None or not len(self.base_nets):
            return None, 0.0
        return engine.net_weight_sum(self.base_nets, self.base_w)

    def _combine(self, mlp_pred, lin):
        pred = mlp_pred.astype(np.float64)
        if lin is not None:
            pred += self.base_ab[0] * lin + self.base_ab[1]
        return pred

    def predict(self, engine):
        """Per-cycle predictions for the trace loaded in `engine`: the base
        term (auxiliary stream) plus the network (its own stream); returns
        (pred, device ms of both)."""
        p, ms = self.mlp.predict_trace(engine, self.feats)
        lin, l_ms = self._base(engine)
        return self._combine(p, lin), ms + l_ms

    def evaluate(self, engine):
        """Per-cycle total power of the trace loaded in `engine`.  The
        network (its own stream), the base term (the engine's auxiliary
        stream) and the judge (the engine's run streams) run concurrently
        from three host threads; all three only read the event store."""
        assert self.feats is not None, "fit() first"
        # the judged tiles, and tile 0: the trace's first cycles (initial
        # values settling) are returned exactly and left out of the error
        tiles = np.union1d(self.judge_tiles(engine.n_cycle), [0]).astype(np.int32)
        t0 = time.perf_counter()
        if self.overlap:
            box = {}

            def run(key, fn):
                try:
                    box[key] = fn()
                except BaseException as e:  # re-raised below
                    box["e"] = e

            ths = [threading.Thread(target=run, args=("j", lambda: engine.power_tiles(self.judge_tile, tiles))),
                   threading.Thread(target=run, args=("b", lambda: self._base(engine)))]
            for th in ths:
                th.start()
            try:
                p, m_ms = self.mlp.predict_trace(engine, self.feats)
            finally:
                for th in ths:
                    th.join()
            if "e" in box:
                raise box["e"]
            cyc, waves, j_ms = box["j"]
            lin, l_ms = box["b"]
            pred, s_ms = self._combine(p, lin), m_ms + l_ms
        else:
            pred, s_ms = self.predict(engine)
            cyc, waves, j_ms = engine.power_tiles(self.judge_tile, tiles)
        yj = cycle_total(waves)
        judged = cyc >= self.skip_cycles
        err = nrmse(pred[cyc[judged]], yj[judged])
        spread = float(yj[judged].std() / yj[judged].mean())
        if self.accepts(err, spread):
            pred[cyc] = yj          # the cycles computed exactly are returned exactly
            wall = (time.perf_counter() - t0) * 1e3
            return SurrogateOutput(pred, "surrogate", err, int(judged.sum()), s_ms, j_ms, 0.0, spread, wall)
        res = engine.power(want=("cycle",))
        return SurrogateOutput(cycle_total(res.cycle_power), "exact", err, int(judged.sum()), s_ms, j_ms, res.total_ms,
                               spread, (time.perf_counter() - t0) * 1e3)
