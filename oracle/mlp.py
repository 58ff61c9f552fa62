"""FP64 restatement of the rank predictor's network (test infrastructure:
the checker of the GPU kernels in oqupy_amd/csrc/pte_rank.hip, never called by
the product).

The network has no counterpart in the reference (the reference always runs
the full SVD of ``split_node_full_svd``, oqupy/backends/node_array.py:262,
285, 541). Parameter layout: W1 [128][24], b1, W2 [128][128], b2, W3, b3,
W4 [128], b4, flattened in that order; hidden activations tanh; loss the mean
squared error. The gradient is back-propagation written out by hand; the tests
pin it against central finite differences.
"""

import numpy as np

FEATURES = 24
WIDTH = 128


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    shapes = [(WIDTH, FEATURES), (WIDTH,), (WIDTH, WIDTH), (WIDTH,),
              (WIDTH, WIDTH), (WIDTH,), (WIDTH,), (1,)]
    out, off = [], 0
    for shp in shapes:
        size = int(np.prod(shp))
        out.append(p[off:off + size].reshape(shp))
        off += size
    assert off == p.size
    return out


def forward(params, x):
    """Outputs and hidden activations for rows x."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.tanh(x @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    h3 = np.tanh(h2 @ w3.T + b3)
    y = h3 @ w4 + b4[0]
    return y, (h1, h2, h3)


def loss_and_grad(params, x, target):
    """Mean squared error and its gradient (flat, the parameter layout)."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(target, dtype=np.float64)
    n = len(t)
    y, (h1, h2, h3) = forward(params, x)
    err = y - t
    loss = float(np.mean(err ** 2))
    dy = 2.0 * err / n
    gw4 = h3.T @ dy
    gb4 = np.array([dy.sum()])
    dz3 = np.outer(dy, w4) * (1.0 - h3 ** 2)
    gw3 = dz3.T @ h2
    gb3 = dz3.sum(axis=0)
    dz2 = (dz3 @ w3) * (1.0 - h2 ** 2)
    gw2 = dz2.T @ h1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ w2) * (1.0 - h1 ** 2)
    gw1 = dz1.T @ x
    gb1 = dz1.sum(axis=0)
    grad = np.concatenate([g.ravel() for g in
                           (gw1, gb1, gw2, gb2, gw3, gb3, gw4, gb4)])
    return loss, grad
