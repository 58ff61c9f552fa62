"""Influence tensors from bath correlation integrals computed on the device.

The reference evaluates every influence tensor (oqupy/tempo.py:969-1020) from
2D correlation integrals that are finite differences of eta(t)
(oqupy/bath_correlations.py:593-664), each eta a QUADPACK integral on the host
(oqupy/bath_correlations.py:518-591): seconds to tens of seconds per process
tensor of a few hundred time steps. Here eta is evaluated on the MI355X for
every grid time t_k = k dt, k = 0 .. dkmax + 1, of one or many baths in one
launch (``pte_eta_batch``: composite Gauss-Legendre, relative accuracy ~1e-13
against tight QUADPACK), and the influence tensors follow from the same
finite differences and exponentials as ``influence_matrix``. Power-law
spectral densities (``PowerLawSD``, zeta >= 1) with exponential, Gaussian or
hard cutoffs; a process tensor with ``add_correlation_time`` needs eta off the
grid and is not served here.

Enabled per computation with ``backend_config={"influence": "device"}``.
"""

from typing import List, Optional, Sequence

import numpy as np
from numpy import ndarray

from oqupy_amd.backends import engine as eng

CUTOFF_TYPES = {"exponential": 0, "gaussian": 1, "hard": 2}


def _bath_row(correlations) -> tuple:
    for attr in ("alpha", "zeta", "cutoff", "cutoff_type", "temperature"):
        if not hasattr(correlations, attr):
            raise NotImplementedError(
                "Device influence tensors need a PowerLawSD bath.")
    if correlations.cutoff_type not in CUTOFF_TYPES:
        raise NotImplementedError(
            f"Cutoff type '{correlations.cutoff_type}' is not supported.")
    if correlations.zeta < 1.0:
        raise NotImplementedError("Device eta needs zeta >= 1.")
    return ((correlations.alpha, correlations.zeta, correlations.cutoff,
             correlations.temperature),
            CUTOFF_TYPES[correlations.cutoff_type])


def eta_values(correlations_list: Sequence, times_list: Sequence[ndarray]
               ) -> List[ndarray]:
    """eta(t) of every bath at its times, all in one device launch."""
    rows, cuts, taus, owner = [], [], [], []
    for b, (corr, times) in enumerate(zip(correlations_list, times_list)):
        row, cut = _bath_row(corr)
        rows.append(row)
        cuts.append(cut)
        t = np.asarray(times, dtype=np.float64).ravel()
        taus.append(t)
        owner.append(np.full(t.size, b, dtype=np.int32))
    params = np.ascontiguousarray(np.array(rows, dtype=np.float64).ravel())
    cut_arr = np.ascontiguousarray(cuts, dtype=np.int32)
    tau_arr = np.ascontiguousarray(np.concatenate(taus))
    own_arr = np.ascontiguousarray(np.concatenate(owner))
    nodes, weights = np.polynomial.legendre.leggauss(16)
    nodes = np.ascontiguousarray(nodes)
    weights = np.ascontiguousarray(weights)
    out = np.zeros(tau_arr.size, dtype=np.complex128)
    eng.require_gpu()
    lib = eng.load_library()
    eng.check(lib.pte_eta_batch(len(rows), eng.dptr(params), eng.iptr(cut_arr),
                                eng.dptr(nodes), eng.dptr(weights),
                                int(tau_arr.size), eng.iptr(own_arr),
                                eng.dptr(tau_arr), eng.dptr(out)))
    res, off = [], 0
    for t in taus:
        res.append(out[off:off + t.size].copy())
        off += t.size
    return res


class DeviceInfluence:
    """``influence(dk)`` of one bath from eta on the time grid, with the
    formulas of ``influence_matrix`` (oqupy/tempo.py:969-1020)."""

    def __init__(self, bath, parameters, dkmax: int,
                 deg_positions: Optional[List[ndarray]] = None,
                 eta: Optional[ndarray] = None) -> None:
        if parameters.add_correlation_time is not None:
            raise NotImplementedError("Device influence tensors do not "
                                      "support add_correlation_time.")
        self._dt = parameters.dt
        self._op_m = bath.coupling_comm
        self._op_p = bath.coupling_acomm
        self._deg = deg_positions
        if eta is None:
            eta = eta_values([bath.correlations],
                             [np.arange(dkmax + 2) * self._dt])[0]
        self.eta = np.asarray(eta)
        self.dkmax = dkmax

    def __call__(self, dk: int) -> Optional[ndarray]:
        if dk < 0:
            return None
        if dk > self.dkmax:
            raise IndexError("dk beyond the precomputed time grid.")
        e = self.eta
        value = e[1] - e[0] if dk == 0 else e[dk + 1] - 2.0 * e[dk] + e[dk - 1]
        op_m, op_p = self._op_m, self._op_p
        if dk == 0:
            infl = np.diag(np.exp(-op_m * (value.real * op_m
                                           + 1j * value.imag * op_p)))
            if self._deg is not None:
                infl = np.diag(infl)[self._deg[0]]
            return infl
        infl = np.exp(-np.outer(value.real * op_m + 1j * value.imag * op_p,
                                op_m))
        if self._deg is not None:
            infl = infl[self._deg[0]][:, self._deg[1]]
        return infl


def device_influences(baths, parameters_list, dkmaxes,
                      deg_positions_list=None) -> List[DeviceInfluence]:
    """DeviceInfluence of many baths with one eta launch for all of them."""
    if deg_positions_list is None:
        deg_positions_list = [None] * len(baths)
    times = [np.arange(k + 2) * p.dt for p, k in zip(parameters_list, dkmaxes)]
    etas = eta_values([b.correlations for b in baths], times)
    return [DeviceInfluence(b, p, k, d, e) for b, p, k, d, e in
            zip(baths, parameters_list, dkmaxes, deg_positions_list, etas)]
