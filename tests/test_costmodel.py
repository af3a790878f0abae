"""Autotuner and learned cost model (MI355X only).

- The device MLP's gradient equals the oracle's (oracle/mlp.py, itself
  pinned by finite differences) within FP32 accuracy.
- Device training lowers the loss; inference matches the oracle forward.
- The front timer runs both variants; per-front variants reach the plan and
  leave the factorization correct.
- What the trained network gains over the method without it: on held-out
  fronts, the variants the model proposes cost less (autotuner-measured
  time) than the default rule's, and its regret against the exhaustive
  autotuner is small.
"""
import numpy as np
import pytest

from sylver_amd import costmodel as CM
from sylver_amd import matrices as M

pytestmark = pytest.mark.gpu


def test_mlp_gradient_vs_oracle(S, gpu):
    from oracle import mlp
    rng = np.random.default_rng(0)
    p = S.mlp_init(7)
    x = rng.normal(0, 1, (300, 8)).astype(np.float32)
    y = rng.normal(0, 1, (300, 2)).astype(np.float32)
    g, loss = S.mlp_grad(p, x, y)
    gr = mlp.grad(p.astype(np.float64), x, y)
    assert abs(loss - mlp.loss(p, x, y)) <= 1e-5 * abs(mlp.loss(p, x, y))
    assert np.abs(g - gr).max() <= 1e-4 * np.abs(gr).max()


def test_mlp_train_and_predict(S, gpu):
    from oracle import mlp
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (2000, 8)).astype(np.float32)
    y = np.stack([np.sin(x[:, 0]) + x[:, 1] ** 2, x[:, 2] * x[:, 3]], 1).astype(np.float32)
    p0 = S.mlp_init(3)
    p, hist = S.mlp_train(p0, x, y, 60, 128, 3e-3, 5)
    assert hist[-1] < 0.2 * hist[0]
    out = S.mlp_predict(p, x[:50])
    ref, _ = mlp.forward(p, x[:50])
    assert np.abs(out - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max())


def test_time_fronts_and_plan_variants(S, gpu):
    t1 = S.time_fronts([200], [100], [1], S.PREC_FP64, 2)
    t0 = S.time_fronts([200], [100], [0], S.PREC_FP64, 2)
    assert t1 > 0 and t0 > 0
    A = M.laplacian_3d(12)
    opts = S.default_options()
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    m, n, _ = CM.tree_fronts(S, akeep)
    b = M.random_rhs(A.n, 2, seed=1)
    for v in (np.ones(len(m), np.int8), np.zeros(len(m), np.int8)):  # all one-workgroup / all tiled
        S.set_front_variants(akeep, v)
        fkeep, finf = S.spldlt_factorize(True, A.val, akeep, options=opts)
        assert finf.flag == 0
        x = np.asfortranarray(b.copy())
        assert S.spldlt_solve(x, akeep, fkeep, options=opts).flag == 0
        assert M.backward_error(A, x, b) <= 1e-15
    S.set_front_variants(akeep, None)


def test_model_gain_over_default_rule(S, gpu):
    """Train on the fronts of a few grids, evaluate on a held-out grid."""
    tuner = CM.Autotuner(S, S.PREC_FP64, reps=2)
    rng = np.random.default_rng(11)

    def fronts(A):
        akeep, _, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=S.default_options())
        m, n, nc = CM.tree_fronts(S, akeep)
        band = CM.in_band(m)
        return m[band], n[band], nc[band]

    xs, ts = [], []
    for A in (M.laplacian_3d(24), M.laplacian_3d(30, 30, 15), M.laplacian_2d(300)):
        m, n, nc = fronts(A)
        xs.append(CM.features(m, n, nc, 0))
        ts.append(tuner.measure(m, n))
    model = CM.CostModel(S, seed=2).fit(np.concatenate(xs), np.concatenate(ts), epochs=200)
    m, n, nc = fronts(M.laplacian_3d(28, 28, 20))
    t = tuner.measure(m, n)
    pick = model.pick_one_wg(CM.features(m, n, nc, 0))
    t_model = np.where(pick, t[:, 0], t[:, 1]).sum()
    t_best = t.min(axis=1).sum()
    t_default = np.where(m <= 256, t[:, 0], t[:, 1]).sum()
    assert t_model <= t_default * 1.001
    assert t_model - t_best <= 0.05 * t_best
    del rng
