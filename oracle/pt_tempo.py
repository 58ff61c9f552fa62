"""PT-TEMPO process-tensor construction on explicit arrays (FP64 oracle).

Restates ``PtTempoBackend`` of ``oqupy/backends/pt_tempo_backend.py``:
``initialize`` (105-190), ``compute_step`` (192-282), ``get_mpo_tensor``
(284-307), and ``SimpleProcessTensor.compute_caps``
(``oqupy/process_tensor.py:380-406``).

Pinning: the golden states ``rho_A``/``rho_C``/``rho_D`` of the reference's
physics tests (through ``oracle.dynamics``), the exact independent-boson
dephasing formula, and the brute-force path sum of ``oracle.path_sum`` at
``epsrel -> 0``.
"""

from typing import Callable, List, Optional

import numpy as np

from oracle import chain


def _create_delta(tensor: np.ndarray, scramble: List[int]) -> np.ndarray:
    """``R[a[s_0], a[s_1], ...] = T[a]`` (oqupy/util.py:30-57)."""
    shape = tensor.shape
    out = np.zeros(tuple(shape[i] for i in scramble), dtype=complex)
    for idx in np.ndindex(*shape):
        out[tuple(idx[i] for i in scramble)] = tensor[idx]
    return out


class PtTempoOracle:
    """FP64 restatement of the PT-TEMPO backend.

    ``influence(dk)`` must return what ``oqupy.tempo.influence_matrix`` returns
    for ``dk`` (a diagonal matrix for ``dk == 0`` without degeneracy maps;
    ``None`` for ``dk < 0`` unless a correlation time is added).
    """

    def __init__(self, dimension: int, influence: Callable[[int], np.ndarray],
                 sum_north: np.ndarray, num_steps: int, dkmax: int,
                 epsrel: float,
                 degeneracy_maps: Optional[List[np.ndarray]] = None) -> None:
        self.dim = dimension
        self.influence = influence
        self.sum_north = np.asarray(sum_north, dtype=complex)
        self.num_steps = num_steps
        self.dkmax = dkmax
        self.epsrel = epsrel
        self.degeneracy_maps = degeneracy_maps
        self.num_infl = min(num_steps, dkmax + 1)
        self.step = None
        self.mps: List[np.ndarray] = []
        self.mpo: List[np.ndarray] = []
        self.bond_log: List[List[int]] = []
        # set to a list to receive (rows, columns, singular values) of every
        # truncated SVD that follows (initialisation sweeps and steps)
        self.spectra: Optional[list] = None

    # -- pt_tempo_backend.py:105-190 ------------------------------------------
    def initialize(self) -> None:
        self.mpo, self.mps = self._build()
        n = len(self.mps)
        chain.sweep_left(self.mps, n - 1, 0, self.epsrel, None, self.spectra)
        chain.sweep_right(self.mps, 0, n - 1, self.epsrel, None, self.spectra)
        self.step = 1

    def load_state(self, step: int, sites: List[np.ndarray]) -> None:
        """Continue from the state another implementation reached after
        ``step`` steps: the MPS sites as given (``[left, physical, right]``)
        and the MPO as the backend holds it at that step."""
        self.mpo, _ = self._build()
        for s in range(2, step + 1):
            self._advance_mpo(s)
        self.mps = [np.array(site, dtype=complex) for site in sites]
        self.step = step

    def _build(self):
        """The initial MPO nodes and MPS sites (before the sweeps)."""
        d = self.dim
        mpo, mps = [], []
        for i in range(self.num_infl):
            infl = self.influence(i)
            if i == 0:
                infl = infl / d
                if self.degeneracy_maps is not None:
                    nmap, wmap = self.degeneracy_maps
                    n_north, n_west = np.max(nmap) + 1, np.max(wmap) + 1
                    node = np.zeros((n_west, d * d, n_north), dtype=complex)
                    site = np.zeros((d * d, n_north), dtype=complex)
                    for i1 in range(d * d):
                        node[wmap[i1], i1, nmap[i1]] = infl[nmap[i1]]
                        site[i1, nmap[i1]] = infl[nmap[i1]] / d
                else:
                    node = _create_delta(infl, [1, 1, 0])
                    site = infl.T / d
                mpo.append(node[None])           # [1, in, out, right]
                mps.append(site[None])           # [1, phys, right]
            elif i == self.num_infl - 1:
                mpo.append(infl[:, None, :, None].astype(complex))
                mps.append(infl[:, :, None].astype(complex))
            else:
                mpo.append(_create_delta(infl, [0, 1, 1, 0]))
                mps.append(_create_delta(infl / d, [0, 1, 0]))
        # (a single node keeps its right leg open: no last-node branch)
        return mpo, mps

    def _advance_mpo(self, step: int) -> bool:
        """The MPO change of ``step`` (pt_tempo_backend.py:229-257); returns
        whether the step is in the end phase."""
        end_phase = step > self.num_steps - self.num_infl + 1
        if end_phase:
            last = self.mpo[-2]
            closed = np.einsum('wpsv,v->wps', last, self.sum_north * self.dim)
            self.mpo = self.mpo[:-2] + [closed[..., None]]
        else:
            infl = self.influence(-step)
            if infl is not None:
                self.mpo[-1] = infl[:, None, :, None].astype(complex)
        return end_phase

    # -- pt_tempo_backend.py:192-282 ------------------------------------------
    def compute_step(self) -> bool:
        self.step += 1
        if not self._advance_mpo(self.step):
            self.mps.append(np.ones((1, 1, 1), dtype=complex))
        log: List[int] = []
        chain.zip_up_left(self.mps, self.mpo, self.epsrel, log, self.spectra)
        chain.sweep_right(self.mps, self.step - 2, len(self.mps) - 1,
                          self.epsrel, log, self.spectra)
        self.bond_log.append(log)
        return self.step < self.num_steps

    def compute(self) -> None:
        """Initialise (if needed) and run every step."""
        if self.step is None:
            self.initialize()
        while self.compute_step():
            pass

    # -- pt_tempo_backend.py:284-307 ------------------------------------------
    def get_mpo_tensor(self, step: int) -> np.ndarray:
        """PT-MPO tensor ``[bond_past, bond_future, d^2]`` of ``step``,
        scaled by the Hilbert dimension."""
        assert len(self.mps) == self.num_steps and step < self.num_steps
        return self.mps[step].transpose(0, 2, 1) * self.dim

    def mpo_tensors(self) -> List[np.ndarray]:
        return [self.get_mpo_tensor(k) for k in range(self.num_steps)]

    def bond_dimensions(self) -> List[int]:
        return [site.shape[2] for site in self.mps[:-1]]


def compute_caps(tensors: List[np.ndarray], dim: int) -> List[np.ndarray]:
    """Caps ``c_k[a] = sum_{b,i} T_k[a, b, i] c_{k+1}[b] tr2[i]`` backward from
    ``[1]``, with ``tr2 = vec(identity)/d`` (oqupy/process_tensor.py:380-406,
    ``_trace_square`` at 55-57)."""
    trace_square = (np.identity(dim) / np.sqrt(float(dim))).reshape(-1) ** 2
    caps = [np.array([1.0], dtype=complex)]
    for ten in reversed(tensors):
        caps.insert(0, np.einsum('abi,b,i->a', ten, caps[0], trace_square))
    return caps
