"""System control operations applied during ``compute_dynamics``.

Restates ``oqupy.control.Control`` (oqupy/control.py:30-198): instantaneous
super-operators applied before ("pre") or after ("post") the recording of
the state at a time step, given by step index (int) or time (float).
"""

from typing import Callable, Optional, Text, Tuple, Union

import numpy as np
from numpy import ndarray

from oqupy_amd.base_api import BaseAPIClass


class Control(BaseAPIClass):
    """A set of system control operations (oqupy/control.py:30-56)."""

    def __init__(self, dimension: int, name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        self._dimension = dimension
        self._step_controls = {"pre": {}, "post": {}}
        self._time_controls = {"pre": {}, "post": {}}
        self._control_times = {"pre": np.array([]), "post": np.array([])}
        super().__init__(name, description)

    @property
    def dimension(self) -> int:
        """Hilbert space dimension of the controlled system."""
        return self._dimension

    def add_single(self, time: Union[int, float], control_operation: ndarray,
                   post: Optional[bool] = False) -> None:
        """Add one operation (a d^2 x d^2 super-operator) at a time step
        (int) or a time (float); several operations at the same step compose
        with the latest applied last (oqupy/control.py:63-107)."""
        key = "post" if post else "pre"
        op = np.asarray(control_operation)
        if isinstance(time, (int, np.integer)) and not isinstance(time, bool):
            controls = self._step_controls[key]
            step = int(time)
            controls[step] = op @ controls[step] if step in controls else op
        elif isinstance(time, float):
            controls = self._time_controls[key]
            if time in self._control_times[key]:
                controls[time] = op @ controls[time]
            else:
                controls[time] = op
                self._control_times[key] = np.sort(
                    np.append(self._control_times[key], time))
        else:
            raise TypeError("Parameter `time` must be either int or float.")

    def add_continuous(self, control_fct: Callable[[ndarray, float], ndarray],
                       post: Optional[bool] = False) -> None:
        """Not provided by the reference either (oqupy/control.py:110-117)."""
        raise NotImplementedError()

    def get_controls(self, step: int, dt: Optional[float] = None,
                     start_time: Optional[float] = 0.0
                     ) -> Tuple[Optional[ndarray], Optional[ndarray]]:
        """(pre, post) super-operators of ``step`` or None for none
        (oqupy/control.py:119-180): time-stamped pre controls, then step pre
        controls; step post controls, then time-stamped post controls."""
        d2 = self._dimension ** 2
        pre = post = None

        def stamped(key):
            times = self._control_times[key]
            if len(times) == 0:
                return []
            hits = np.round((times - start_time) / dt) == step
            return [self._time_controls[key][t] for t in times[hits]]

        for op in stamped("pre"):
            pre = op @ (np.identity(d2) if pre is None else pre)
        if step in self._step_controls["pre"]:
            pre = self._step_controls["pre"][step] @ \
                (np.identity(d2) if pre is None else pre)
        if step in self._step_controls["post"]:
            post = self._step_controls["post"][step] @ np.identity(d2)
        for op in stamped("post"):
            post = op @ (np.identity(d2) if post is None else post)
        return pre, post
