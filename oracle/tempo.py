"""TEMPO step on explicit arrays (FP64 oracle).

Restates ``BaseTempoBackend.initialize_mps_mpo`` and ``compute_system_step``
of ``oqupy/backends/tempo_backend.py:379-575`` (and ``TempoBackend``,
578-626). Site arrays are ``[left, physical, right]``; influence nodes are
``[west, north, south, east]``.

Pinning: golden states ``rho_A``/``rho_C``/``rho_D`` of the reference physics
tests, the exact dephasing formula, basis-rotation invariance, and the
brute-force path sum at ``epsrel -> 0``.
"""

from typing import Callable, List, Optional

import numpy as np

from oracle import chain
from oracle.pt_tempo import _create_delta


def _left_right_super(left: np.ndarray, right: np.ndarray) -> np.ndarray:
    return np.kron(left, right.T)


class TempoOracle:
    """FP64 restatement of ``TempoBackend``."""

    def __init__(self, initial_state: np.ndarray,
                 influence: Callable[[int], np.ndarray],
                 unitary_transform: np.ndarray,
                 propagators: Callable[[int], tuple],
                 sum_north: np.ndarray, sum_west: np.ndarray,
                 dkmax: Optional[int], epsrel: float,
                 degeneracy_maps: Optional[List[np.ndarray]] = None,
                 dim: Optional[int] = None) -> None:
        self.initial_state = np.array(initial_state, dtype=complex).reshape(-1)
        self.influence = influence
        self.unitary = np.asarray(unitary_transform)
        self.propagators = propagators
        self.sum_north = np.asarray(sum_north, dtype=complex)
        self.sum_west = np.asarray(sum_west, dtype=complex)
        self.dkmax = dkmax
        self.epsrel = epsrel
        self.degeneracy_maps = degeneracy_maps
        self.dim = dim
        self.step = None
        self.bond_log: List[List[int]] = []

    def _four_legs(self, infl: np.ndarray) -> np.ndarray:
        return _create_delta(infl, [1, 0, 0, 1])

    # -- tempo_backend.py:379-437 ---------------------------------------------
    def initialize(self):
        u = self.unitary
        super_u = _left_right_super(u, u.conjugate().T)
        super_u_dagg = _left_right_super(u.conjugate().T, u)
        count = 1 if self.dkmax is None else self.dkmax + 1
        nodes = []
        for i in range(count):
            infl = self.influence(i)
            if i == 0:
                if self.degeneracy_maps is not None:
                    nmap, wmap = self.degeneracy_maps
                    d2 = self.dim ** 2
                    node = np.zeros((np.max(wmap) + 1, d2, np.max(nmap) + 1,
                                     d2), dtype=complex)
                    for i1 in range(d2):
                        node[wmap[i1], i1, nmap[i1], i1] = infl[nmap[i1]]
                else:
                    node = self._four_legs(infl)
                    node = np.einsum('wnse,nm,fe->wmsf', node, super_u_dagg,
                                     super_u)
            else:
                node = self._four_legs(infl)
            nodes.append(node)
        self.mpo = list(reversed(nodes))     # [dk=K, ..., dk=1, dk=0]
        self.mps = [self.initial_state.reshape(1, -1, 1)]
        self.step = 0
        return 0, self.initial_state.copy()

    # -- tempo_backend.py:439-575 ---------------------------------------------
    def compute_system_step(self, step: int, prop_1, prop_2) -> np.ndarray:
        if self.dkmax is None:
            window = list(self.mpo)
            self.mpo = [self._four_legs(self.influence(len(window)))] \
                + self.mpo
        elif step <= self.dkmax:
            window = list(self.mpo[-step:])
        else:
            window = list(self.mpo)
            infl = self.influence(self.dkmax - step)
            if infl is not None:
                window[0] = self._four_legs(infl)
        # close the west leg of the oldest node with sum_west (line 519)
        window[0] = np.einsum('wnse,w->nse', window[0],
                              self.sum_west)[None]
        # first half-step on the newest site's physical leg (lines 521-529)
        self.mps[-1] = np.einsum('lpr,qp->lqr', self.mps[-1], prop_1)
        # trace out the oldest site once the window is full (lines 531-537)
        if len(self.mps) != len(window):
            traced = np.einsum('lpr,p->lr', self.mps[0], self.sum_north)
            self.mps[1] = np.einsum('lr,rqs->lqs', traced, self.mps[1])
            self.mps = self.mps[1:]
        log: List[int] = []
        chain.zip_up_right(self.mps, window, self.epsrel, log)
        # the last site keeps the east leg of the dk=0 node as its right
        # bond; sweep right to left (lines 549-553)
        n = len(self.mps)
        chain.sweep_left(self.mps, n - 1, 0, self.epsrel, log)
        # join the second half-step: its row index is the east leg
        # (lines 555-558): site[e, p] = prop_2.T[e, p]
        self.mps.append(prop_2.T.reshape(prop_2.shape[1], -1, 1)
                        .astype(complex))
        self.bond_log.append(log)
        # read out the state: sum every physical leg but the newest one
        vec = np.ones((1,), dtype=complex)
        for site in self.mps[:-1]:
            vec = np.einsum('l,lpr,p->r', vec, site, self.sum_north)
        state = np.einsum('l,lpr->p', vec, self.mps[-1])
        return state

    def compute_step(self):
        self.step += 1
        prop_1, prop_2 = self.propagators(self.step - 1)
        state = self.compute_system_step(self.step, prop_1, prop_2)
        return self.step, state
