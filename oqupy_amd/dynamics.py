"""Time-sorted density matrices and their expectation values.

Restates ``BaseDynamics``/``Dynamics`` of ``oqupy/dynamics.py:26-184`` and the
input helpers at 300-369: states are kept sorted by time, and
``expectations(O)`` returns ``Tr(O rho(t))`` for every recorded time.
"""

from bisect import bisect
from copy import copy
from typing import List, Optional, Text, Tuple

import numpy as np
from numpy import ndarray

from oqupy_amd.base_api import BaseAPIClass
from oqupy_amd.config import NpDtype, NpDtypeReal


class BaseDynamics(BaseAPIClass):
    """Times and states of an evolution (oqupy/dynamics.py:26-131)."""

    def __init__(self, name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        super().__init__(name, description)
        self._times = []
        self._states = []
        self._shape = None

    def __str__(self) -> Text:
        ret = super().__str__() + f"  length        = {len(self)} timesteps \n"
        if len(self) > 0:
            ret += f"  min time      = {np.min(self._times)} \n"
            ret += f"  max time      = {np.max(self._times)} \n"
        return ret

    def __len__(self) -> int:
        return len(self._times)

    @property
    def times(self) -> ndarray:
        """Times of the dynamics."""
        return np.array(self._times, dtype=NpDtypeReal)

    @property
    def states(self) -> ndarray:
        """States of the dynamics."""
        return np.array(self._states, dtype=NpDtype)

    @property
    def shape(self):
        """Shape of the states."""
        return copy(self._shape)

    def expectations(self, operator: Optional[ndarray] = None,
                     real: Optional[bool] = False) -> Tuple[ndarray, ndarray]:
        """``(times, Tr(O rho(t)))``; the trace of rho if ``operator`` is
        None (oqupy/dynamics.py:78-131)."""
        if len(self) == 0:
            return None, None
        if operator is None:
            op = np.identity(self._shape[0], dtype=NpDtype)
        else:
            try:
                op = np.array(operator, dtype=NpDtype)
            except Exception as err:
                raise AssertionError("Argument `operator` must be ndarray.") \
                    from err
            assert op.shape == tuple(self._shape), (
                "Argument `operator` must have the same shape as the states. "
                f"Has shape {op.shape}, but should be {self._shape}.")
        values = np.array([np.trace(op @ state) for state in self._states])
        times = np.array(self._times)
        return times, (np.real(values) if real else values)


class Dynamics(BaseDynamics):
    """A time evolution of a density matrix (oqupy/dynamics.py:133-184)."""

    def __init__(self, times: Optional[List[float]] = None,
                 states: Optional[List[ndarray]] = None,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        super().__init__(name, description)
        times = [] if times is None else times
        states = [] if states is None else states
        assert isinstance(times, list), "Argument `times` must be a list."
        assert isinstance(states, list), "Argument `states` must be a list."
        assert len(times) == len(states), \
            "Lists `times` and `states` must have the same length."
        for time, state in zip(times, states):
            self.add(time, state)

    def add(self, time: float, state: ndarray) -> None:
        """Insert ``state`` at ``time``, keeping the times sorted."""
        try:
            t = float(time)
        except Exception as err:
            raise AssertionError("Argument `time` must be float.") from err
        try:
            rho = np.array(state, dtype=NpDtype)
        except Exception as err:
            raise AssertionError("Argument `state` must be ndarray.") from err
        if self._shape is not None:
            assert rho.shape == tuple(self._shape), (
                "Appended state doesn't have the same shape as previous "
                f"states ({rho.shape}, but should be {self._shape})")
        assert rho.ndim == 2 and rho.shape[0] == rho.shape[1], (
            f"State must be a square matrix. But the dimensions are "
            f"{rho.shape}.")
        index = bisect(self._times, t)
        self._times.insert(index, t)
        self._states.insert(index, rho)
        if self._shape is None:
            self._shape = rho.shape

    @property
    def shape(self):
        """The shape of the states."""
        return np.copy(self._shape)


class MeanFieldDynamics(BaseAPIClass):
    """The time evolution of the systems of a mean-field system and of their
    coherent field (oqupy/dynamics.py:186-296): one ``Dynamics`` per system
    plus the field at every time."""

    def __init__(self, times: Optional[List[float]] = None,
                 system_states_list: Optional[List[List[ndarray]]] = None,
                 fields: Optional[List[complex]] = None,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        super().__init__(name, description)
        self._times = []
        self._fields = []
        self._system_dynamics = []
        times = [] if times is None else times
        fields = [] if fields is None else fields
        system_states_list = [] if system_states_list is None \
            else system_states_list
        assert isinstance(times, list), "Argument `times` must be a list."
        assert isinstance(fields, list), "Argument `fields` must be a list."
        assert isinstance(system_states_list, list), \
            "Argument `system_states_list` must be a list."
        assert len(times) == len(fields) == len(system_states_list), \
            "Lists `times`, `system_states_list` and `fields` must have the " \
            "same length."
        for time, states, field in zip(times, system_states_list, fields):
            self.add(time, states, field)

    def add(self, time: float, system_states: List[ndarray],
            field: complex) -> None:
        """Insert every system's state and the field at ``time``."""
        try:
            t = float(time)
        except Exception as err:
            raise AssertionError("Argument `time` must be float.") from err
        try:
            fld = complex(field)
        except Exception as err:
            raise AssertionError("Argument `field` must be complex.") from err
        if not self._system_dynamics:
            self._system_dynamics = [Dynamics(name=f"system {i}")
                                     for i in range(len(system_states))]
        else:
            assert len(system_states) == len(self._system_dynamics), (
                f"Number of states to add ({len(system_states)}) does not "
                "match number of Dynamics objects in MeanFieldDynamics "
                f"({len(self._system_dynamics)})")
        index = bisect(self._times, t)
        self._times.insert(index, t)
        self._fields.insert(index, fld)
        for dyn, state in zip(self._system_dynamics, system_states):
            dyn.add(t, state)

    def __len__(self) -> int:
        return len(self._times)

    @property
    def times(self) -> ndarray:
        """Times of the dynamics."""
        return np.array(self._times, dtype=NpDtypeReal)

    @property
    def fields(self) -> ndarray:
        """The field at every time."""
        return np.array(self._fields, dtype=NpDtype)

    def field_expectations(self):
        """``(times, fields)``, or ``(None, None)`` when empty."""
        if len(self) == 0:
            return None, None
        return self.times, self.fields

    @property
    def system_dynamics(self) -> List[Dynamics]:
        """One Dynamics per system."""
        return self._system_dynamics
