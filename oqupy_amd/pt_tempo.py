"""PT-TEMPO frontend: process-tensor construction.

Restates ``PtTempo`` and ``pt_tempo_compute`` of ``oqupy/pt_tempo.py:64-375``
(input checks, number of steps, influence callable, basis transforms of a
non-diagonal coupling, resumable step loop). The steps run in the MI355X
engine (``oqupy_amd.backends.pt_tempo_backend.PtTempoBackend``).
"""

import ctypes
from copy import copy
from typing import Dict, List, Optional, Text, Union

import numpy as np
from numpy import ndarray

from oqupy_amd.backends.engine import parse_backend_config as eng_config
from oqupy_amd.base_api import BaseAPIClass
from oqupy_amd.bath import Bath
from oqupy_amd.config import PT_TEMPO_BACKEND_CONFIG
from oqupy_amd.operators import left_right_super
from oqupy_amd.process_tensor import BaseProcessTensor, SimpleProcessTensor
from oqupy_amd.tempo import (TempoParameters, degeneracy_positions,
                             influence_matrix)
from oqupy_amd.util import get_progress


class PtTempo(BaseAPIClass):
    """A PT-TEMPO computation of the process tensor of one bath
    (oqupy/pt_tempo.py:64-306)."""

    def __init__(self, bath: Bath, start_time: float, end_time: float,
                 parameters: TempoParameters,
                 unique: Optional[bool] = False,
                 process_tensor_file: Optional[Union[Text, bool]] = None,
                 overwrite: Optional[bool] = False,
                 backend_config: Optional[Dict] = None,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        assert isinstance(bath, Bath), \
            "Argument 'bath' must be an instance of Bath."
        self._bath = bath
        self._dimension = bath.dimension
        self._correlations = bath.correlations
        super().__init__(name, description)
        try:
            self._start_time = float(start_time)
        except Exception as err:
            raise AssertionError("Start time must be a float.") from err
        try:
            self._end_time = float(end_time)
        except Exception as err:
            raise AssertionError("End time must be a float.") from err
        assert isinstance(parameters, TempoParameters), \
            "Argument 'parameters' must be an instance of TempoParameters."
        self._parameters = parameters
        assert isinstance(unique, bool), "Argument 'unique' must be a boolean."
        self._unique = unique
        if process_tensor_file:
            raise NotImplementedError(
                "File-backed (HDF5) process tensors are not provided.")
        self._init_simple_process_tensor()
        self._backend_config = PT_TEMPO_BACKEND_CONFIG \
            if backend_config is None else backend_config
        num_steps = int((end_time - self._start_time) / self._parameters.dt)
        assert num_steps >= 2, (
            "Parameter `end_time` must be more than two times steps larger "
            "than the parameter `start_time`!")
        self._num_steps = num_steps
        # influence tensors from eta evaluated on the device (built at first
        # use, or handed in by pt_tempo_compute_batch for many baths at once)
        self._device_influence = None
        self._use_device_influence = \
            eng_config(self._backend_config)["influence"] == "device"
        self._backend_instance = None
        self._init_pt_tempo_backend()

    def _init_simple_process_tensor(self) -> None:
        """oqupy/pt_tempo.py:157-174."""
        unitary = self._bath.unitary_transform
        if not np.allclose(unitary, np.identity(self._dimension)):
            transform_in = left_right_super(unitary.conjugate().T, unitary).T
            transform_out = left_right_super(unitary, unitary.conjugate().T).T
        else:
            transform_in = transform_out = None
        self._process_tensor = SimpleProcessTensor(
            hilbert_space_dimension=self._dimension, dt=self._parameters.dt,
            transform_in=transform_in, transform_out=transform_out,
            name=self.name, description=self.description)

    def _init_pt_tempo_backend(self) -> None:
        """oqupy/pt_tempo.py:202-228, with the MI355X engine."""
        from oqupy_amd.backends.pt_tempo_backend import PtTempoBackend
        if self._unique:
            sum_north = np.ones(np.max(self._bath.north_degeneracy_map) + 1)
            sum_west = np.ones(np.max(self._bath.west_degeneracy_map) + 1)
            maps = [self._bath.north_degeneracy_map,
                    self._bath.west_degeneracy_map]
        else:
            sum_north = np.ones(self._dimension ** 2)
            sum_west = np.ones(self._dimension ** 2)
            maps = None
        dkmax = self._parameters.dkmax
        if dkmax is None:
            dkmax = self._num_steps
        config = dict(self._backend_config)
        if config.get("rank_predictor") is not None \
                or config.get("rank_trace", 0):
            # bath features of the rank predictor
            config.setdefault("rank_bath", (
                float(getattr(self._correlations, "alpha", 0.0)),
                float(getattr(self._correlations, "temperature", 0.0))))
        self._backend_instance = PtTempoBackend(
            dimension=self._dimension, influence=self._influence,
            process_tensor=self._process_tensor, sum_north=sum_north,
            sum_west=sum_west, num_steps=self._num_steps, dkmax=dkmax,
            epsrel=self._parameters.epsrel, config=config,
            degeneracy_maps=maps)

    def _influence(self, dk: int) -> ndarray:
        if self._use_device_influence:
            if self._device_influence is None:
                self._device_influence = self._make_device_influence()
            return self._device_influence(dk)
        positions = degeneracy_positions(self._bath) if self._unique else None
        return influence_matrix(dk, parameters=self._parameters,
                                correlations=self._correlations,
                                coupling_acomm=self._bath.coupling_acomm,
                                coupling_comm=self._bath.coupling_comm,
                                deg_positions=positions)

    def _make_device_influence(self, eta=None):
        from oqupy_amd.influence import DeviceInfluence
        dkmax = self._parameters.dkmax
        if dkmax is None:
            dkmax = self._num_steps
        positions = degeneracy_positions(self._bath) if self._unique else None
        return DeviceInfluence(self._bath, self._parameters, dkmax, positions,
                               eta)

    @property
    def dimension(self) -> int:
        """Hilbert space dimension."""
        return copy(self._dimension)

    @property
    def num_steps(self) -> int:
        """Number of time steps of the process tensor."""
        return self._num_steps

    @property
    def backend(self):
        """The engine backend (bond dimensions, step statistics)."""
        return self._backend_instance

    def compute(self, progress_type: Optional[Text] = None) -> None:
        """Run (or continue) the computation (oqupy/pt_tempo.py:258-278)."""
        if self._backend_instance.step is None:
            self._backend_instance.initialize()
        progress = get_progress(progress_type)
        with progress(self._backend_instance.num_steps,
                      "--> PT-TEMPO computation:") as prog:
            while self._backend_instance.compute_step():
                prog.update(self._backend_instance.step)
            prog.update(self._backend_instance.step)

    def get_process_tensor(self, progress_type: Optional[Text] = None
                           ) -> BaseProcessTensor:
        """The process tensor, computed if needed (oqupy/pt_tempo.py:280-306)."""
        if self._backend_instance.step is None or \
                self._backend_instance.step < self._backend_instance.num_steps:
            self.compute(progress_type=progress_type)
        if len(self._process_tensor) < self._backend_instance.num_steps:
            self._backend_instance.update_process_tensor()
        return self._process_tensor


def pt_tempo_compute(bath: Bath, start_time: float, end_time: float,
                     parameters: Optional[TempoParameters] = None,
                     unique: Optional[bool] = False,
                     tolerance: Optional[float] = None,
                     process_tensor_file: Optional[Union[Text, bool]] = None,
                     overwrite: Optional[bool] = False,
                     backend_config: Optional[Dict] = None,
                     progress_type: Optional[Text] = None,
                     name: Optional[Text] = None,
                     description: Optional[Text] = None) -> BaseProcessTensor:
    """Create a ``PtTempo``, compute and return the process tensor
    (oqupy/pt_tempo.py:309-375). ``parameters`` must be given (parameter
    estimation is not part of this package)."""
    if parameters is None:
        raise NotImplementedError(
            "TEMPO parameter estimation is not provided; pass `parameters`.")
    ptt = PtTempo(bath, start_time, end_time, parameters, unique,
                  process_tensor_file, overwrite, backend_config, name,
                  description)
    ptt.compute(progress_type=progress_type)
    return ptt.get_process_tensor()


def pt_tempo_compute_batch(baths: List[Bath],
                           start_time: Union[float, List[float]],
                           end_time: Union[float, List[float]],
                           parameters: Union[TempoParameters,
                                             List[TempoParameters]],
                           unique: Optional[bool] = False,
                           backend_config: Optional[Dict] = None,
                           names: Optional[List[Text]] = None
                           ) -> List[BaseProcessTensor]:
    """Process tensors of many baths computed together: the engines of all
    instances are stepped in lockstep and every truncated SVD of one step is
    served by batched device launches (no counterpart in the reference; each
    result equals ``pt_tempo_compute(bath_i, start_i, end_i, parameters_i)``,
    oqupy/pt_tempo.py:309-375). ``start_time``, ``end_time`` and
    ``parameters`` are either one value for all baths or one per bath.
    Parameters with ``add_correlation_time`` are not supported here."""
    from oqupy_amd.backends import engine as eng
    count = len(baths)
    if count == 0:
        return []

    def per_bath(value, what):
        if isinstance(value, (list, tuple)):
            if len(value) != count:
                raise ValueError(f"`{what}` needs one entry per bath.")
            return list(value)
        return [value] * count

    starts = per_bath(start_time, "start_time")
    ends = per_bath(end_time, "end_time")
    params = per_bath(parameters, "parameters")
    names = per_bath(names, "names") if names is not None else [None] * count
    for p in params:
        if p is None:
            raise NotImplementedError(
                "TEMPO parameter estimation is not provided; pass "
                "`parameters`.")
        if p.add_correlation_time is not None:
            raise NotImplementedError(
                "Batched PT-TEMPO does not support `add_correlation_time`.")
    ptts = [PtTempo(b, s, e, p, unique, None, False, backend_config, n)
            for b, s, e, p, n in zip(baths, starts, ends, params, names)]
    if ptts[0]._use_device_influence:
        # eta of every bath on its time grid in one device launch
        from oqupy_amd.influence import eta_values
        dkmaxes = [p.dkmax if p.dkmax is not None else t.num_steps
                   for p, t in zip(params, ptts)]
        etas = eta_values([b.correlations for b in baths],
                          [np.arange(k + 2) * p.dt
                           for p, k in zip(params, dkmaxes)])
        for ptt, eta in zip(ptts, etas):
            ptt._device_influence = ptt._make_device_influence(eta)
    backends = [p.backend for p in ptts]
    for b in backends:
        b.initialize()
    lib = eng.load_library()
    handles = (ctypes.c_void_p * count)(*[b.handle.value for b in backends])
    batch = ctypes.c_void_p()
    eng.check(lib.pte_pt_batch_create(handles, count, ctypes.byref(batch)))
    try:
        more = np.ones(count, dtype=np.int32)
        steps = max(p.num_steps for p in ptts)
        while more.any():
            eng.check(lib.pte_pt_batch_run(batch, steps, eng.iptr(more)))
    finally:
        lib.pte_pt_batch_destroy(batch)
    for b in backends:
        b.sync_step()
    return [p.get_process_tensor() for p in ptts]
