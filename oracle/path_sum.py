"""Brute-force path sum of the discretised influence functional.

For a diagonal coupling operator the TEMPO network of
``oqupy/backends/tempo_backend.py:379-575``, contracted without truncation,
equals the sum over Liouville-space paths ``a_0 .. a_{N-1}`` (one index per
time point, taken after the first half-step propagator):

    rho_N[i] = sum_a (P1 rho_0)[a_0] prod_{m=1}^{N-1} (P1 P2)[a_m, a_{m-1}]
               * P2[i, a_{N-1}]
               * prod_{m=0}^{N-1} prod_{k=0}^{min(m, K)} I_k[a_{m-k}, a_m]

where ``I_0`` is the diagonal dk=0 influence (``I_0[a, a] = D_0[a]``) and, for
``m >= K`` with an added correlation time, ``I_K`` of the pair
``(a_{m-K}, a_m)`` is the rectangle influence ``I(K - m - 1)``
(tempo_backend.py:502-516). This restatement is written from the network
structure, independent of the chain code, and costs ``4^N`` terms: it is meant
for N <= 8. It is the exact (epsrel -> 0) limit that both chain oracles and the
MI355X engine must reproduce.
"""

from typing import Callable

import numpy as np


def path_sum(initial_state: np.ndarray, prop_1: np.ndarray,
             prop_2: np.ndarray, influence: Callable[[int], np.ndarray],
             num_steps: int, dkmax: int) -> np.ndarray:
    """``rho(t_N)`` as a ``d x d`` matrix (time-independent propagators)."""
    d2 = prop_1.shape[0]
    n = num_steps
    d0 = np.diag(influence(0)) if np.ndim(influence(0)) == 2 \
        else influence(0)
    paths = np.array(np.unravel_index(np.arange(d2 ** n), (d2,) * n)).T
    # propagator chain
    head = (prop_1 @ initial_state.reshape(-1))[paths[:, 0]]
    mid = prop_1 @ prop_2
    amp = head.astype(complex)
    for m in range(1, n):
        amp = amp * mid[paths[:, m], paths[:, m - 1]]
    # influence factors
    for m in range(n):
        amp = amp * d0[paths[:, m]]
        for k in range(1, min(m, dkmax) + 1):
            infl = influence(k)
            if k == dkmax and m >= dkmax:
                rect = influence(dkmax - m - 1)
                if rect is not None:
                    infl = rect
            amp = amp * infl[paths[:, m - k], paths[:, m]]
    rho = np.zeros(d2, dtype=complex)
    for i in range(d2):
        rho[i] = np.sum(amp * prop_2[i, paths[:, n - 1]])
    dim = int(round(np.sqrt(d2)))
    return rho.reshape(dim, dim)
