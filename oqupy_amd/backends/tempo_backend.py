"""TEMPO backend running on the MI355X engine.

Keeps the interface of the reference's ``TempoBackend``
(oqupy/backends/tempo_backend.py:578-626): ``initialize() -> (0, state)``,
``compute_step() -> (step, state)`` and ``step``. Influences and propagators
are evaluated on the host where the reference calls them and passed to the
engine as arrays; the tensor chain stays in device memory.
"""

import ctypes
from copy import copy
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np
from numpy import ndarray

from oqupy_amd.backends import engine as eng


class TempoBackend:
    """TEMPO on the MI355X engine."""

    def __init__(self, initial_state: ndarray,
                 influence: Callable[[int], ndarray],
                 unitary_transform: ndarray,
                 propagators: Callable[[int], Tuple[ndarray, ndarray]],
                 sum_north: ndarray, sum_west: ndarray, dkmax: Optional[int],
                 epsrel: float, config: Optional[Dict] = None,
                 degeneracy_maps: Optional[List[ndarray]] = None,
                 dim: Optional[int] = None) -> None:
        self._initial_state = np.asarray(initial_state).reshape(-1)
        self._influence = influence
        self._unitary = np.asarray(unitary_transform)
        self._propagators = propagators
        self._sum_north = np.asarray(sum_north)
        self._sum_west = np.asarray(sum_west)
        self._dkmax = dkmax
        self._epsrel = epsrel
        self._config = eng.parse_backend_config(config)
        self._degeneracy_maps = degeneracy_maps
        self._dim = dim if dim is not None else self._unitary.shape[0]
        self._step = None
        self._state = None
        self._handle = None
        self._lib = eng.load_library()

    @property
    def step(self) -> Optional[int]:
        """The current step (None before ``initialize``)."""
        return self._step

    def _set_influence(self, dk: int, infl: ndarray) -> None:
        arr = eng.as_c128(infl)
        cols = arr.shape[1] if arr.ndim == 2 else 0
        eng.check(self._lib.pte_tempo_set_influence(
            self._handle, dk, eng.dptr(arr), arr.shape[0], cols))

    def initialize(self) -> Tuple[int, ndarray]:
        """Build the MPO and the one-site MPS (tempo_backend.py:379-437,
        610-617)."""
        eng.require_gpu()
        cfg = eng._TempoConfig(
            dimension=self._dim,
            dkmax=-1 if self._dkmax is None else self._dkmax,
            epsrel=self._epsrel,
                            precision=eng.PRECISIONS[self._config["precision"]],
            chi_cap=self._config["chi_cap"],
            n_north=int(self._sum_north.size),
            n_west=int(self._sum_west.size))
        handle = ctypes.c_void_p()
        eng.check(self._lib.pte_tempo_create(ctypes.byref(cfg),
                                             ctypes.byref(handle)))
        self._handle = handle
        eng.track(self)
        count = 1 if self._dkmax is None else self._dkmax + 1
        for dk in range(count):
            self._set_influence(dk, self._influence(dk))
        if self._degeneracy_maps is not None:
            nmap = np.ascontiguousarray(self._degeneracy_maps[0],
                                        dtype=np.int32)
            wmap = np.ascontiguousarray(self._degeneracy_maps[1],
                                        dtype=np.int32)
            eng.check(self._lib.pte_tempo_set_degeneracy_maps(
                self._handle, eng.iptr(nmap), eng.iptr(wmap), int(nmap.size)))
        rho0 = eng.as_c128(self._initial_state)
        unitary = eng.as_c128(self._unitary)
        eng.check(self._lib.pte_tempo_initialize(self._handle, eng.dptr(rho0),
                                                 eng.dptr(unitary)))
        self._step = 0
        self._state = self._initial_state.astype(np.complex128)
        return self._step, copy(self._state)

    def compute_step(self) -> Tuple[int, ndarray]:
        """One TEMPO step (tempo_backend.py:619-626)."""
        step = self._step + 1
        prop_1, prop_2 = self._propagators(step - 1)
        return step, self.compute_system_step(step, prop_1, prop_2)

    def compute_system_step(self, step: int, prop_1: ndarray,
                            prop_2: ndarray) -> ndarray:
        """Step ``step`` with the given half-step propagators; returns the
        new state vector (BaseTempoBackend.compute_system_step,
        tempo_backend.py:439-575)."""
        assert step == self._step + 1, "steps must follow each other"
        extra = None
        rows = cols = 0
        if self._dkmax is None:
            self._set_influence(step, self._influence(step))
        elif step > self._dkmax:
            infl = self._influence(self._dkmax - step)
            if infl is not None:
                extra = eng.as_c128(infl)
                rows, cols = extra.shape
        p1 = eng.as_c128(prop_1)
        p2 = eng.as_c128(prop_2)
        state = np.empty(self._dim ** 2, dtype=np.complex128)
        eng.check(self._lib.pte_tempo_step(self._handle, eng.dptr(p1),
                                           eng.dptr(p2), eng.dptr(extra),
                                           rows, cols, eng.dptr(state)))
        self._step = step
        self._state = state
        return copy(self._state)

    def bond_dimensions(self) -> List[int]:
        """Bond dimensions of the device-resident MPS."""
        n = ctypes.c_int()
        eng.check(self._lib.pte_tempo_bond_dims(self._handle, None, 0,
                                                ctypes.byref(n)))
        out = np.zeros(max(n.value - 1, 1), dtype=np.int32)
        eng.check(self._lib.pte_tempo_bond_dims(
            self._handle, eng.iptr(out), int(out.size), ctypes.byref(n)))
        return [int(b) for b in out[:max(n.value - 1, 0)]]

    def release(self) -> None:
        """Destroy the engine and free its device memory (idempotent)."""
        handle = getattr(self, "_handle", None)
        self._handle = None
        if handle is not None and handle.value and not eng.finalizing():
            self._lib.pte_tempo_destroy(handle)

    def __del__(self) -> None:
        try:
            self.release()
        except Exception:  # interpreter shutdown
            pass


class MeanFieldTempoBackend:
    """Several TEMPO computations and a coherent field stepped together
    (oqupy/backends/tempo_backend.py:629-773): one engine per system; the
    field equation of motion and the field-dependent propagators are
    evaluated on the host between the steps, where the reference evaluates
    them."""

    def __init__(self, initial_state_list: List[ndarray],
                 initial_field: complex,
                 influence_list: List[Callable[[int], ndarray]],
                 unitary_transform_list: List[ndarray],
                 propagators_list: List[Callable[[int, complex, complex],
                                                 Tuple[ndarray, ndarray]]],
                 compute_field: Callable, compute_field_derivative: Callable,
                 sum_north_list: List[ndarray], sum_west_list: List[ndarray],
                 dkmax: Optional[int], epsrel: float,
                 config: Optional[Dict] = None,
                 degeneracy_maps_list: Optional[List] = None,
                 dim_list: Optional[List[int]] = None) -> None:
        count = len(initial_state_list)
        if degeneracy_maps_list is None:
            degeneracy_maps_list = [None] * count
        if dim_list is None:
            dim_list = [None] * count
        self._states = [np.asarray(s).reshape(-1) for s in initial_state_list]
        self._field = complex(initial_field)
        self._propagators_list = propagators_list
        self._compute_field = compute_field
        self._compute_field_derivative = compute_field_derivative
        self._step = None
        # the propagators of each system are passed per step
        self._backends = [
            TempoBackend(state, influence, unitary, None, north, west, dkmax,
                         epsrel, config, maps, dim)
            for state, influence, unitary, north, west, maps, dim in zip(
                self._states, influence_list, unitary_transform_list,
                sum_north_list, sum_west_list, degeneracy_maps_list,
                dim_list)]

    @property
    def step(self) -> Optional[int]:
        """The current step (None before ``initialize``)."""
        return self._step

    def initialize(self) -> Tuple[int, List[ndarray], complex]:
        """Initialise every system's network (tempo_backend.py:749-754)."""
        for backend in self._backends:
            backend.initialize()
        self._step = 0
        return self._step, [copy(s) for s in self._states], self._field

    def compute_step(self) -> Tuple[int, List[ndarray], complex]:
        """One step of every system with the field and its derivative at the
        start of the step, then the field's own step
        (tempo_backend.py:756-773)."""
        step = self._step
        states = [copy(s) for s in self._states]
        field = self._field
        derivative = self._compute_field_derivative(step, states, field)
        props = [prop(step, field, derivative)
                 for prop in self._propagators_list]
        next_states = [backend.compute_system_step(step + 1, *pair)
                       for backend, pair in zip(self._backends, props)]
        self._field = self._compute_field(step, states, field, next_states)
        self._states = next_states
        self._step = step + 1
        return self._step, [copy(s) for s in self._states], self._field

    def release(self) -> None:
        """Destroy every system's engine."""
        for backend in self._backends:
            backend.release()
