"""CPU checks of the learned rank form's host side: the FP64 network oracle
against finite differences, the parameter layout, the trace-record helpers
and the configuration checks."""

import numpy as np
import pytest

from oracle import mlp as mlp_oracle
from oqupy_amd import rank_predictor as rp
from oqupy_amd.backends import engine


def _case(seed=0, n=40):
    rng = np.random.default_rng(seed)
    params = rp.RankPredictor.initial_params(seed).astype(np.float64)
    params += 0.05 * rng.normal(size=params.size)     # non-zero biases too
    x = rng.normal(size=(n, rp.FEATURES))
    t = rng.normal(size=n) * 0.3
    return params, x, t


def test_param_layout():
    assert rp.param_count() == 36353
    assert rp.RankPredictor.initial_params().size == rp.param_count()
    assert len(mlp_oracle.unpack(np.zeros(rp.param_count()))) == 8


def test_oracle_gradient_matches_finite_differences():
    params, x, t = _case()
    loss, grad = mlp_oracle.loss_and_grad(params, x, t)
    rng = np.random.default_rng(1)
    # every block of the layout, plus random entries
    offs = np.cumsum([0] + [a.size for a in mlp_oracle.unpack(params)])
    picks = set(int(o) for o in offs[:-1]) | set(int(o) - 1 for o in offs[1:])
    picks |= set(rng.integers(0, params.size, size=60).tolist())
    h = 1e-6
    for p in sorted(picks):
        e = np.zeros_like(params)
        e[p] = h
        lp, _ = mlp_oracle.loss_and_grad(params + e, x, t)
        lm, _ = mlp_oracle.loss_and_grad(params - e, x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[p]) <= 1e-7 + 1e-5 * abs(fd), p
    assert loss == pytest.approx(np.mean((mlp_oracle.forward(params, x)[0]
                                          - t) ** 2))


def test_oracle_gradient_detects_a_wrong_activation_derivative():
    """A plausible slip (tanh' taken as 1 - h instead of 1 - h^2) fails the
    finite-difference check."""
    params, x, t = _case(2)
    _, grad = mlp_oracle.loss_and_grad(params, x, t)
    w1, b1, w2, b2, w3, b3, w4, b4 = mlp_oracle.unpack(params)
    _, (h1, h2, h3) = mlp_oracle.forward(params, x)
    y = h3 @ w4 + b4[0]
    dy = 2.0 * (y - t) / len(t)
    wrong_gb3 = (np.outer(dy, w4) * (1.0 - h3)).sum(axis=0)
    off = sum(a.size for a in (w1, b1, w2, b2, w3))
    assert np.abs(wrong_gb3 - grad[off:off + rp.WIDTH]).max() > 1e-4


def test_records_helpers():
    rec = np.zeros((5, rp.RECORD), dtype=np.float32)
    rec[:, 19] = [0.0, 0.5, 0.5, 0.0, 0.25]
    rec[:, rp.FEATURES + 1] = [3, 4, 5, 6, 7]
    use = rp.usable(rec)
    assert len(use) == 3
    feats, target, ranks = rp.split_records(use)
    assert feats.shape == (3, rp.FEATURES) and list(ranks) == [4, 5, 7]
    with pytest.raises(ValueError):
        rp.split_records(np.zeros((2, 5)))


def test_rank_config_checks():
    cfg = engine.parse_backend_config({"rank_predictor": "previous",
                                       "rank_trace": 100})
    assert cfg["rank_predictor"] == "previous" and cfg["rank_trace"] == 100
    assert engine.parse_backend_config(None)["rank_predictor"] is None
    with pytest.raises(ValueError):
        engine.parse_backend_config({"rank_predictor": "mlp"})
    with pytest.raises(ValueError):
        engine.parse_backend_config({"rank_predictor": "previous",
                                     "precision": "c64"})
    with pytest.raises(ValueError):
        engine.parse_backend_config({"rank_theta": 0.0})
    pred = rp.RankPredictor(seed=3)
    assert engine.parse_backend_config(
        {"rank_predictor": pred})["rank_predictor"] is pred
    with pytest.raises(ValueError):
        rp.RankPredictor(np.zeros(10))


def test_save_load_roundtrip(tmp_path):
    pred = rp.RankPredictor(seed=4)
    pred.loss_history = [1.0, 0.5]
    path = str(tmp_path / "pred.npz")
    pred.save(path)
    back = rp.RankPredictor.load(path)
    assert np.array_equal(back.params, pred.params)
    assert back.loss_history == [1.0, 0.5]
