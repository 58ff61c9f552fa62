"""Rank predictor of the learned truncation-rank form.

Every truncated SVD of PT-TEMPO's zip-up and sweep starts with a pivoted QR
that, in the exact form, runs until the residual is at the level of machine
precision. The learned form caps it: a small MLP (24 features -> 3 x 128 tanh
-> 1, FP32) predicts how many steps ("need") the pivoted QR takes before its
residual falls below theta * epsrel * |beta_0| (|beta_0| the largest column
norm, a lower bound of s0), the QR is stopped at the prediction times
``1 + rank_margin`` plus ``rank_oversample`` steps, and the Jacobi SVD runs on
the smaller factor; the residual then counts as discarded weight in the
epsrel cut. When the capped factorisation's residual is above the threshold
the QR resumes where it stopped and runs to the exact stop, which also checks
the predictor.

The engine records training data while it runs (``backend_config=
{"rank_trace": capacity}``), the network is trained on the GPU by the
engine's own forward, back-propagation and Adam kernels
(``pte_rank_mlp_train``), and a trained predictor is handed back through
``backend_config={"rank_predictor": predictor}``. Two forms without the
network stand beside it: ``"previous"`` caps at the bond's previous step count
plus the oversampling, ``"adaptive"`` stops the QR at the threshold itself.

Features of one truncation (``pte_rank_feature_count()`` = 24): step / N,
distance of the site from the newest one / dkmax, kind (0 zip-up, 1 sweep),
the bond's last 16 kept singular values at its previous truncation as
log10(s / s0) / 8, log2 of the steps it needed then / 8, alpha, temperature,
log10(epsrel) / 10 and log2(min(m, n)) / 8. Record target: log2(need) / 8; the
network is trained on (and predicts) its change from the previous truncation,
log2(need / previous need) / 8, which carries over to memory lengths and bond
dimensions outside the training runs.
"""

from typing import Optional

import numpy as np

from oqupy_amd.backends import engine as eng

FEATURES = 24
WIDTH = 128
RECORD = 28


def param_count() -> int:
    """Number of FP32 parameters of the network."""
    return WIDTH * FEATURES + WIDTH + 2 * (WIDTH * WIDTH + WIDTH) + WIDTH + 1


def split_records(records: np.ndarray):
    """(features, targets, need) of trace records from ``rank_trace()``."""
    rec = np.asarray(records, dtype=np.float32)
    if rec.ndim != 2 or rec.shape[1] != RECORD:
        raise ValueError(f"trace records must be (n, {RECORD}).")
    return (np.ascontiguousarray(rec[:, :FEATURES]),
            np.ascontiguousarray(rec[:, FEATURES]),
            rec[:, FEATURES + 1].astype(np.int64))


def usable(records: np.ndarray) -> np.ndarray:
    """Records whose bond had a previous truncation (the only ones the engine
    predicts for)."""
    rec = np.asarray(records, dtype=np.float32)
    return rec[rec[:, 19] > 0.0]


class RankPredictor:
    """The MLP's parameters and its training / inference calls."""

    def __init__(self, params: Optional[np.ndarray] = None,
                 seed: int = 0) -> None:
        if params is None:
            params = self.initial_params(seed)
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != param_count():
            raise ValueError("wrong number of parameters.")
        self.params = params
        self.loss_history = []

    @staticmethod
    def initial_params(seed: int = 0) -> np.ndarray:
        """Glorot-uniform weights, zero biases."""
        rng = np.random.default_rng(seed)
        parts = []
        for fan_in, fan_out in ((FEATURES, WIDTH), (WIDTH, WIDTH),
                                (WIDTH, WIDTH), (WIDTH, 1)):
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            parts.append(rng.uniform(-lim, lim, size=fan_out * fan_in))
            parts.append(np.zeros(fan_out))
        return np.concatenate(parts).astype(np.float32)

    def fit(self, records: np.ndarray, epochs: int = 40, batch: int = 256,
            lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        """Train on trace records (Adam, mean squared error on the change of
        log2(steps)/8 from the bond's previous truncation); returns the mean
        loss of every epoch."""
        feats, target, _ = split_records(usable(records))
        if len(target) == 0:
            raise ValueError("no usable training records.")
        # the network learns the change from the bond's previous step count
        target = np.ascontiguousarray(target - feats[:, 19])
        eng.require_gpu()
        lib = eng.load_library()
        losses = np.zeros(epochs, dtype=np.float32)
        params = self.params.copy()
        eng.check(lib.pte_rank_mlp_train(
            eng.fptr(params), eng.fptr(feats), eng.fptr(target), len(target),
            int(epochs), int(batch), float(lr), int(seed) & 0xffffffff,
            eng.fptr(losses)))
        self.params = params
        self.loss_history.extend(losses.tolist())
        return losses

    def predict(self, features: np.ndarray) -> np.ndarray:
        """Network outputs for rows of features, on the GPU: the predicted
        change of log2(steps) / 8 from the bond's previous truncation."""
        feats = np.ascontiguousarray(features, dtype=np.float32)
        if feats.ndim != 2 or feats.shape[1] != FEATURES:
            raise ValueError(f"features must be (n, {FEATURES}).")
        eng.require_gpu()
        lib = eng.load_library()
        out = np.zeros(len(feats), dtype=np.float32)
        if len(feats):
            eng.check(lib.pte_rank_mlp_predict(eng.fptr(self.params),
                                               eng.fptr(feats), len(feats),
                                               eng.fptr(out)))
        return out

    def predict_steps(self, features: np.ndarray) -> np.ndarray:
        """Predicted pivoted-QR step counts (not rounded): the network's
        change added to the bond's previous log2(steps) / 8."""
        feats = np.asarray(features, dtype=np.float32)
        out = self.predict(feats).astype(np.float64) + feats[:, 19]
        return np.exp2(8.0 * out)

    def gradient(self, features: np.ndarray, targets: np.ndarray):
        """(loss, gradient) of the mean squared error, on the GPU."""
        feats = np.ascontiguousarray(features, dtype=np.float32)
        tgt = np.ascontiguousarray(targets, dtype=np.float32)
        eng.require_gpu()
        lib = eng.load_library()
        grad = np.zeros(param_count(), dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        eng.check(lib.pte_rank_mlp_grad(eng.fptr(self.params), eng.fptr(feats),
                                        eng.fptr(tgt), len(tgt),
                                        eng.fptr(grad), eng.fptr(loss)))
        return float(loss[0]), grad

    def save(self, path: str) -> None:
        """Write the parameters (numpy .npz, no pickled objects)."""
        np.savez(path, params=self.params,
                 loss_history=np.asarray(self.loss_history, dtype=np.float32))

    @classmethod
    def load(cls, path: str) -> "RankPredictor":
        """Read parameters written by ``save``."""
        with np.load(path, allow_pickle=False) as data:
            pred = cls(data["params"])
            pred.loss_history = data["loss_history"].tolist()
        return pred
