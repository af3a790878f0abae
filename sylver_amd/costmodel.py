"""Autotuner and learned cost model for the per-front kernel variant of the
posdef factorization: the one-workgroup kernel (fronts factored whole by
one workgroup, batched per level) or the tiled large-front path (diagonal
block, TRSM and Schur-update launches spread over the GPU).

- `Autotuner` times every front shape of interest under both variants on
  synthetic SPD fronts (`sylver_amd_time_fronts`, device time of the
  factorization launches alone) and keeps the faster one: the exhaustive
  search and the judge of the model's regret.
- `CostModel` is the library's 8 -> 64 -> 64 -> 2 FP32 MLP, trained on the
  device (`sylver_amd_mlp_train`) on the autotuner's timings, predicting
  log2 of the time of both variants from eight shape features; its choice
  is the variant with the smaller prediction.
- `choose_variants` turns either into the per-front variant array of
  `set_front_variants`. Fronts outside [M_MIN, M_MAX] rows keep the default
  rule (the one-workgroup kernel below, the tiled path above), where one
  variant wins by a wide margin.
"""
from __future__ import annotations

import math

import numpy as np

M_MIN, M_MAX = 32, 2048  # rows of the fronts whose variant is chosen


def node_flops(m, n):
    """Reference flop model of one front: sum over its pivots r = m-n+1..m of r^2."""
    m = np.asarray(m, np.float64)
    q = m - np.asarray(n, np.float64)
    s = lambda k: k * (k + 1) * (2 * k + 1) / 6.0  # noqa: E731
    return s(m) - s(q)


def features(m, n, nchild, tier):
    """Eight inputs per front: log2 m, log2 n, log2(m-n+1), n/m, log2 flops,
    log2(children+1), tier (0 FP64, 1 FP32), (m mod 128)/128."""
    m = np.asarray(m, np.float64)
    n = np.asarray(n, np.float64)
    nchild = np.asarray(nchild, np.float64)
    tier = np.broadcast_to(np.asarray(tier, np.float64), m.shape)
    return np.stack([np.log2(m), np.log2(n), np.log2(m - n + 1), n / m, np.log2(node_flops(m, n)),
                     np.log2(nchild + 1), tier, np.mod(m, 128) / 128.0], axis=1).astype(np.float32)


def tree_fronts(S, akeep):
    """(m, n, nchild) of every front of an analysis."""
    sym = S.akeep_symbolic(akeep)
    m = np.diff(sym["rptr"]).astype(np.int64)
    n = np.diff(sym["sptr"]).astype(np.int64)
    par = sym["sparent"] - 1
    nn = len(m)
    nchild = np.bincount(par[par < nn], minlength=nn) if nn else np.zeros(0, np.int64)
    return m, n, nchild


def bucket(m, n):
    """Shape bucket of a front: m rounded up to 16 (to 64 above 512), n to
    the same step (at most m)."""
    step = 16 if m <= 512 else 64
    mb = int(math.ceil(m / step) * step)
    nb = min(mb, int(math.ceil(n / step) * step))
    return mb, max(nb, 1)


class Autotuner:
    """Exhaustive timing of both variants per shape bucket (cached)."""

    def __init__(self, S, precision, reps=3):
        self.S, self.precision, self.reps = S, precision, reps
        self.cache = {}

    def times(self, mb, nb):
        key = (mb, nb)
        if key not in self.cache:
            t1 = self.S.time_fronts([mb], [nb], [1], self.precision, self.reps)
            t0 = self.S.time_fronts([mb], [nb], [0], self.precision, self.reps)
            self.cache[key] = (t1, t0)  # (one workgroup, tiled)
        return self.cache[key]

    def measure(self, m, n):
        """(t_one_wg, t_tiled) in ms for every front (through its bucket)."""
        out = np.zeros((len(m), 2))
        for i, (mi, ni) in enumerate(zip(m, n)):
            out[i] = self.times(*bucket(int(mi), int(ni)))
        return out


def in_band(m):
    m = np.asarray(m)
    return (m >= M_MIN) & (m <= M_MAX)


def choose_variants(m, pick_one_wg):
    """Variant array for set_front_variants: the choice inside the band,
    -1 (default rule) outside."""
    v = np.full(len(m), -1, np.int8)
    band = in_band(m)
    v[band] = np.where(np.asarray(pick_one_wg)[band], 1, 0)
    return v


class CostModel:
    """The device-trained MLP: predicts log2(ms) of (one workgroup, tiled)."""

    def __init__(self, S, seed=0):
        self.S = S
        self.params = S.mlp_init(seed)
        self.loss = None

    def fit(self, x, t, epochs=300, batch=256, lr=3e-3, seed=1, max_records=200_000, max_steps=60_000):
        """Train on (features, times); at most max_records records (a seeded
        subsample) and about max_steps optimizer steps."""
        x = np.asarray(x, np.float32)
        y = np.log2(np.maximum(np.asarray(t, np.float64), 1e-4)).astype(np.float32)
        if len(x) > max_records:
            keep = np.random.default_rng(seed).choice(len(x), max_records, replace=False)
            x, y = x[keep], y[keep]
        steps_per_epoch = max(1, -(-len(x) // batch))
        epochs = int(max(5, min(epochs, max_steps // steps_per_epoch)))
        self.params, self.loss = self.S.mlp_train(self.params, x, y, epochs, batch, lr, seed)
        return self

    def predict(self, x):
        return self.S.mlp_predict(self.params, x)

    def pick_one_wg(self, x):
        p = self.predict(x)
        return p[:, 0] < p[:, 1]
