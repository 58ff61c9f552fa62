"""TEMPO: parameters, influence tensors and the time-stepping frontend.

Restates the parts of ``oqupy/tempo.py`` on the process-tensor path:
``TempoParameters`` (50-216), ``Tempo`` (277-491), ``MeanFieldTempo``
(665-960), ``influence_matrix`` (969-1020), ``tempo_compute`` (1142-1209) and
``_parameter_memory_input_parse`` (1254-1276). The step itself runs in the
MI355X engine (``oqupy_amd.backends.tempo_backend.TempoBackend``, one per
system for the mean-field form).
"""

from copy import copy
from typing import Dict, List, Optional, Text, Union

import numpy as np
from numpy import ndarray

from oqupy_amd.base_api import BaseAPIClass
from oqupy_amd.bath import Bath
from oqupy_amd.bath_correlations import BaseCorrelations
from oqupy_amd.config import (INTEGRATE_EPSREL, SUBDIV_LIMIT,
                              TEMPO_BACKEND_CONFIG)
from oqupy_amd.dynamics import Dynamics, MeanFieldDynamics
from oqupy_amd.system import (BaseSystem, MeanFieldSystem, System,
                              TimeDependentSystem,
                              TimeDependentSystemWithField)
from oqupy_amd.util import (check_convert, check_isinstance, check_true,
                            get_progress)


class TempoParameters(BaseAPIClass):
    """Parameters of a TEMPO or PT-TEMPO computation
    (oqupy/tempo.py:50-216).

    ``dt`` is the time step, ``epsrel`` the relative SVD truncation error (also
    the quadrature tolerance of the eta integrals), and the memory is given by
    at most one of ``tcut`` (a time) or ``dkmax`` (a number of steps).
    """

    def __init__(self, dt: float, epsrel: float,
                 tcut: Optional[float] = None, dkmax: Optional[int] = None,
                 add_correlation_time: Optional[float] = None,
                 subdiv_limit: Optional[int] = SUBDIV_LIMIT,
                 liouvillian_epsrel: Optional[float] = INTEGRATE_EPSREL,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        try:
            tmp_dt = float(dt)
        except Exception as err:
            raise TypeError("Argument 'dt' must be float.") from err
        if tmp_dt <= 0.0:
            raise ValueError("Argument 'dt' must be positive.")
        self._dt = tmp_dt

        try:
            tmp_epsrel = float(epsrel)
        except Exception as err:
            raise TypeError("Argument 'epsrel' must be float.") from err
        if tmp_epsrel <= 0.0:
            raise ValueError("Argument 'epsrel' must be positive.")
        self._epsrel = tmp_epsrel

        self._tcut, self._dkmax = _parameter_memory_input_parse(tcut, dkmax,
                                                                dt)
        try:
            tau = None if add_correlation_time is None \
                else float(add_correlation_time)
        except Exception as err:
            raise TypeError("Argument 'add_correlation_time' must be float "
                            "or None.") from err
        if tau is not None and tau < 0:
            raise ValueError(
                "Argument 'add_correlation_time' must be non-negative.")
        self._add_correlation_time = tau

        try:
            limit = None if subdiv_limit is None else int(subdiv_limit)
        except Exception as err:
            raise TypeError("Argument 'subdiv_limit' must be int or "
                            "None.") from err
        if limit is not None and limit < 0:
            raise ValueError(
                "Argument 'subdiv_limit' must be non-negative or None.")
        self._subdiv_limit = limit

        try:
            liou_eps = float(liouvillian_epsrel)
        except Exception as err:
            raise TypeError("Argument 'liouvillian_epsrel' must be "
                            "float.") from err
        if liou_eps <= 0.0:
            raise ValueError("Argument 'liouvillian_epsrel' must be "
                             "positive.")
        self._liouvillian_epsrel = liou_eps
        super().__init__(name, description)

    def __str__(self) -> Text:
        return (super().__str__()
                + f"  dt                   = {self.dt} \n"
                + f"  tcut [dkmax]         = {self.tcut} [{self.dkmax}] \n"
                + f"  epsrel               = {self.epsrel} \n"
                + f"  add_correlation_time = {self.add_correlation_time} \n")

    @property
    def dt(self) -> float:
        """Length of a time step."""
        return self._dt

    @property
    def epsrel(self) -> float:
        """Maximal relative error of the singular value truncation."""
        return self._epsrel

    @property
    def tcut(self) -> Optional[float]:
        """Length of the non-Markovian memory."""
        return self._tcut

    @property
    def dkmax(self) -> Optional[int]:
        """Number of time steps in the non-Markovian memory."""
        return self._dkmax

    @property
    def add_correlation_time(self) -> Optional[float]:
        """Extra correlation time folded into the last influence."""
        return self._add_correlation_time

    @property
    def subdiv_limit(self) -> Optional[int]:
        """Subdivision limit when integrating a time-dependent Liouvillian."""
        return self._subdiv_limit

    @property
    def liouvillian_epsrel(self) -> float:
        """Relative tolerance when integrating a time-dependent Liouvillian."""
        return self._liouvillian_epsrel


def _parameter_memory_input_parse(tcut, dkmax, dt):
    """``(tcut, dkmax)`` from at most one of them (oqupy/tempo.py:1254-1276):
    ``tcut = dkmax * dt`` or ``dkmax = ceil(round(tcut / dt))``."""
    if tcut is not None and dkmax is not None:
        raise AssertionError(
            "Only one of 'tcut', 'dkmax' should be specified.")
    if dkmax is not None:
        if not isinstance(dkmax, int):
            raise TypeError("Argument 'dkmax' must be int or None.")
        if dkmax < 0:
            raise ValueError("Argument 'dkmax' must be non-negative.")
        return dkmax * dt, dkmax
    if tcut is not None:
        try:
            tmp_tcut = float(tcut)
        except Exception as err:
            raise TypeError("Argument 'tcut' must be float or None.") from err
        if tmp_tcut < 0:
            raise ValueError("Argument 'tcut' must be non-negative.")
        return tmp_tcut, int(np.ceil(np.round(tcut / dt)))
    return None, None


def influence_matrix(dk: int, parameters: TempoParameters,
                     correlations: BaseCorrelations, coupling_acomm: ndarray,
                     coupling_comm: ndarray,
                     deg_positions: Optional[List[ndarray]] = None):
    """Influence tensor of time-step distance ``dk`` (oqupy/tempo.py:969-1020).

    * ``dk == 0``: ``diag(exp(-m * (Re(eta) m + i Im(eta) p)))`` from the
      upper-triangle integral over the first step (a d^2 x d^2 diagonal
      matrix, or its diagonal restricted to ``deg_positions[0]``);
    * ``dk > 0``: ``exp(-outer(Re(eta) m + i Im(eta) p, m))`` from the square
      integral at ``dk * dt`` (rows: the older point, columns: the newer);
    * ``dk < 0``: the rectangle integral from ``dkmax * dt`` when a correlation
      time is added, else ``None``;

    with ``m``/``p`` the diagonals of the coupling's commutator/
    anti-commutator super-operators. The 2D integrals use
    ``epsrel = parameters.epsrel``, as the reference does.
    """
    dt = parameters.dt
    dkmax = parameters.dkmax
    if dk == 0:
        time_1, time_2, shape = 0.0, None, "upper-triangle"
    elif dk < 0:
        if parameters.add_correlation_time is None:
            return None
        time_1 = float(dkmax) * dt
        time_2 = float(dkmax) * dt + np.min(
            [float(-dk) * dt, 1.0 * dt + parameters.add_correlation_time])
        shape = "rectangle"
    else:
        time_1, time_2, shape = float(dk) * dt, None, "square"

    eta = correlations.correlation_2d_integral(
        delta=dt, time_1=time_1, time_2=time_2, shape=shape,
        epsrel=parameters.epsrel)
    op_p = coupling_acomm
    op_m = coupling_comm
    if dk == 0:
        infl = np.diag(np.exp(-op_m * (eta.real * op_m + 1j * eta.imag * op_p)))
        if deg_positions is not None:
            infl = np.diag(infl)[deg_positions[0]]
        return infl
    infl = np.exp(-np.outer(eta.real * op_m + 1j * eta.imag * op_p, op_m))
    if deg_positions is not None:
        north_pos, west_pos = deg_positions
        infl = infl[north_pos][:, west_pos]
    return infl


def degeneracy_positions(bath: Bath) -> List[ndarray]:
    """First index of every distinct north/west class
    (oqupy/tempo.py:362-370, pt_tempo.py:232-240)."""
    north = bath.north_degeneracy_map
    west = bath.west_degeneracy_map
    return [np.array([np.where(north == i)[0][0]
                      for i in range(np.max(north) + 1)]),
            np.array([np.where(west == i)[0][0]
                      for i in range(np.max(west) + 1)])]


def _check_time(end_time) -> float:
    try:
        return float(end_time)
    except Exception as err:
        raise TypeError("End time must be a float.") from err


def _tempo_physical_input_parse(system, initial_state, bath,
                                with_field=False):
    """oqupy/tempo.py:1278-1300; ``with_field`` takes a system of a
    mean-field system."""
    if with_field:
        check_isinstance(system, TimeDependentSystemWithField, "system")
    else:
        check_isinstance(system, (System, TimeDependentSystem), "system")
    hs_dim = system.dimension
    check_isinstance(initial_state, ndarray, "initial_state")
    check_true(initial_state.shape == (hs_dim, hs_dim),
               "Initial sate must be a square matrix of "
               f"dimension {hs_dim}x{hs_dim}.")
    check_isinstance(bath, Bath, 'bath')
    check_true(bath.dimension == hs_dim,
               "Hilbertspace dimensions are unequal: "
               f"system ({hs_dim}), and bath coupling ({bath.dimension}).")
    return system, initial_state, bath, hs_dim


class Tempo(BaseAPIClass):
    """The TEMPO tensor network of a system coupled to one bath
    (oqupy/tempo.py:277-491); each step runs on the MI355X engine."""

    def __init__(self, system: Union[System, TimeDependentSystem],
                 bath: Bath, parameters: TempoParameters,
                 initial_state: ndarray, start_time: float,
                 unique: Optional[bool] = False,
                 backend_config: Optional[Dict] = None,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        super().__init__(name, description)
        assert isinstance(bath, Bath), \
            "Argument 'bath' must be an instance of Bath."
        self._system, self._initial_state, self._bath, self._dimension = \
            _tempo_physical_input_parse(system, initial_state, bath)
        self._correlations = self._bath.correlations
        assert isinstance(parameters, TempoParameters), \
            "Argument 'parameters' must be an instance of TempoParameters."
        self._parameters = parameters
        try:
            self._start_time = float(start_time)
        except Exception as err:
            raise TypeError("Start time must be a float.") from err
        assert isinstance(unique, bool), "Argument 'unique' must be a boolean."
        self._unique = unique
        self._backend_config = TEMPO_BACKEND_CONFIG if backend_config is None \
            else backend_config
        self._dynamics = None
        self._backend_instance = None
        self._prepare_backend()

    def _influence(self, dk: int) -> ndarray:
        positions = degeneracy_positions(self._bath) if self._unique else None
        return influence_matrix(dk, parameters=self._parameters,
                                correlations=self._correlations,
                                coupling_acomm=self._bath.coupling_acomm,
                                coupling_comm=self._bath.coupling_comm,
                                deg_positions=positions)

    def _time(self, step: int) -> float:
        return self._start_time + float(step) * self._parameters.dt

    def _get_num_step(self, start_step: int, end_time: float) -> int:
        end_step = int((end_time - self._start_time) / self._parameters.dt)
        return max(0, end_step - start_step)

    @property
    def dimension(self) -> int:
        """Hilbert space dimension."""
        return copy(self._dimension)

    def _prepare_backend(self) -> None:
        """oqupy/tempo.py:400-435, with the MI355X engine as the backend."""
        from oqupy_amd.backends.tempo_backend import TempoBackend
        dim = self._dimension
        propagators = self._system.get_propagators(
            self._parameters.dt, self._start_time,
            self._parameters.subdiv_limit, self._parameters.liouvillian_epsrel)
        if self._unique:
            sum_north = np.ones(np.max(self._bath.north_degeneracy_map) + 1)
            sum_west = np.ones(np.max(self._bath.west_degeneracy_map) + 1)
            maps = [self._bath.north_degeneracy_map,
                    self._bath.west_degeneracy_map]
        else:
            sum_north = np.ones(dim ** 2)
            sum_west = np.ones(dim ** 2)
            maps = None
        self._backend_instance = TempoBackend(
            self._initial_state.reshape(dim ** 2), self._influence,
            self._bath.unitary_transform, propagators, sum_north, sum_west,
            self._parameters.dkmax, self._parameters.epsrel,
            config=self._backend_config, degeneracy_maps=maps, dim=dim)

    def compute(self, end_time: float, progress_type: Text = None
                ) -> Dynamics:
        """Propagate (or continue) up to ``end_time``
        (oqupy/tempo.py:444-486)."""
        end = _check_time(end_time)
        dim = self._dimension
        if self._backend_instance.step is None:
            step, state = self._backend_instance.initialize()
            self._dynamics = Dynamics(
                description=f"computed from '{self.name}' tempo")
            self._dynamics.add(self._time(step), state.reshape(dim, dim))
        num_step = self._get_num_step(self._backend_instance.step, end)
        with get_progress(progress_type)(num_step, "--> TEMPO computation:") \
                as prog:
            for i in range(num_step):
                prog.update(i)
                step, state = self._backend_instance.compute_step()
                self._dynamics.add(self._time(step), state.reshape(dim, dim))
            prog.update(num_step)
        return self._dynamics

    def get_dynamics(self) -> Dynamics:
        """The dynamics computed so far."""
        return self._dynamics


def tempo_compute(system: BaseSystem, bath: Bath, initial_state: ndarray,
                  start_time: float, end_time: float,
                  parameters: Optional[TempoParameters] = None,
                  tolerance: Optional[float] = None,
                  unique: Optional[bool] = False,
                  backend_config: Optional[Dict] = None,
                  progress_type: Optional[Text] = None,
                  name: Optional[Text] = None,
                  description: Optional[Text] = None) -> Dynamics:
    """Create a ``Tempo`` and compute up to ``end_time``
    (oqupy/tempo.py:1142-1209). Parameter estimation
    (``guess_tempo_parameters``) is not part of this package: ``parameters``
    must be given."""
    if parameters is None:
        raise NotImplementedError(
            "TEMPO parameter estimation is not provided; pass `parameters`.")
    tempo = Tempo(system, bath, parameters, initial_state, start_time, unique,
                  backend_config, name, description)
    tempo.compute(end_time, progress_type=progress_type)
    return tempo.get_dynamics()


class MeanFieldTempo(BaseAPIClass):
    """TEMPO for every system of a ``MeanFieldSystem`` together with the
    evolution of their coherent field (oqupy/tempo.py:665-960). Each system
    has its own bath and TEMPO network on the MI355X engine; per step the
    field derivative at the step's start sets the field-dependent
    propagators of every system, and the field itself advances by the
    second-order Runge-Kutta rule over the step."""

    def __init__(self, mean_field_system: MeanFieldSystem,
                 bath_list: List[Bath], parameters: TempoParameters,
                 initial_state_list: List[ndarray], initial_field: complex,
                 start_time: Optional[float] = 0.0,
                 unique: Optional[bool] = False,
                 backend_config: Optional[Dict] = None,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        assert isinstance(mean_field_system, MeanFieldSystem), \
            "Argument 'mean_field_system' must be an instance of " \
            "MeanFieldSystem."
        self._mean_field_system = mean_field_system
        self._dynamics = None
        self._backend_config = TEMPO_BACKEND_CONFIG if backend_config is None \
            else backend_config
        assert isinstance(parameters, TempoParameters), \
            "Argument 'parameters' must be an instance of TempoParameters."
        self._parameters = parameters
        assert isinstance(unique, bool), "Argument 'unique' must be a boolean."
        self._unique = unique
        super().__init__(name, description)
        assert isinstance(bath_list, list), \
            "bath_list must be a list of Bath objects"
        assert isinstance(initial_state_list, list), \
            "initial_state_list must be a list of state matrices"
        systems = mean_field_system.system_list
        assert len(systems) == len(bath_list) == len(initial_state_list), (
            f"The lengths of bath_list ({len(bath_list)}) and "
            f"initial_state_list ({len(initial_state_list)}) must match the "
            f"number ({len(systems)}) of systems in mean_field_system.")
        parsed = [_tempo_physical_input_parse(sys_, rho, bath, with_field=True)
                  for sys_, rho, bath in zip(systems, initial_state_list,
                                             bath_list)]
        self._systems = [p[0] for p in parsed]
        self._initial_states = [p[1] for p in parsed]
        self._baths = [p[2] for p in parsed]
        self._dims = [p[3] for p in parsed]
        self._initial_field = check_convert(initial_field, complex,
                                            "initial_field")
        self._start_time = check_convert(start_time, float, "start_time")
        self._prepare_backend()

    def _influence(self, bath: Bath):
        positions = degeneracy_positions(bath) if self._unique else None

        def influence(dk: int) -> ndarray:
            return influence_matrix(dk, parameters=self._parameters,
                                    correlations=bath.correlations,
                                    coupling_acomm=bath.coupling_acomm,
                                    coupling_comm=bath.coupling_comm,
                                    deg_positions=positions)
        return influence

    def _prepare_backend(self) -> None:
        """oqupy/tempo.py:770-830, with one engine per system."""
        from oqupy_amd.backends.tempo_backend import MeanFieldTempoBackend
        par = self._parameters
        props = [s.get_propagators(par.dt, self._start_time,
                                   par.subdiv_limit, par.liouvillian_epsrel)
                 for s in self._systems]
        if self._unique:
            norths = [np.ones(np.max(b.north_degeneracy_map) + 1)
                      for b in self._baths]
            wests = [np.ones(np.max(b.west_degeneracy_map) + 1)
                     for b in self._baths]
            maps = [[b.north_degeneracy_map, b.west_degeneracy_map]
                    for b in self._baths]
        else:
            norths = [np.ones(d ** 2) for d in self._dims]
            wests = [np.ones(d ** 2) for d in self._dims]
            maps = [None] * len(self._baths)
        self._backend_instance = MeanFieldTempoBackend(
            [rho.reshape(-1) for rho in self._initial_states],
            self._initial_field, [self._influence(b) for b in self._baths],
            [b.unitary_transform for b in self._baths], props,
            self._compute_field, self._compute_field_derivative, norths,
            wests, par.dkmax, par.epsrel, config=self._backend_config,
            degeneracy_maps_list=maps, dim_list=self._dims)

    def _time(self, step: int) -> float:
        return self._start_time + float(step) * self._parameters.dt

    def _matrices(self, states: List[ndarray]) -> List[ndarray]:
        return [np.asarray(s).reshape(d, d) for s, d in zip(states, self._dims)]

    def _compute_field(self, step: int, state_list: List[ndarray],
                       field: complex, next_state_list: List[ndarray]
                       ) -> complex:
        """The field at step + 1 by second-order Runge-Kutta (Heun) with the
        states at both ends of the step (oqupy/tempo.py:867-882)."""
        dt = self._parameters.dt
        t = self._time(step)
        eom = self._mean_field_system.field_eom
        slope_0 = eom(t, self._matrices(state_list), field)
        slope_1 = eom(t + dt, self._matrices(next_state_list),
                      field + slope_0 * dt)
        return field + dt * (slope_0 + slope_1) / 2

    def _compute_field_derivative(self, step: int,
                                  state_list: List[ndarray],
                                  field: complex) -> complex:
        """d field / dt at ``step`` (oqupy/tempo.py:884-892)."""
        return self._mean_field_system.field_eom(
            self._time(step), self._matrices(state_list), field)

    def compute(self, end_time: float, progress_type: Text = None
                ) -> MeanFieldDynamics:
        """Propagate (or continue) every system and the field up to
        ``end_time`` (oqupy/tempo.py:894-940)."""
        end = _check_time(end_time)
        if self._backend_instance.step is None:
            step, states, field = self._backend_instance.initialize()
            self._dynamics = MeanFieldDynamics(
                description=f"computed from '{self.name}' MeanFieldTempo")
            self._dynamics.add(self._time(step), self._matrices(states), field)
        start = self._backend_instance.step
        num_step = max(0, int((end - self._start_time) / self._parameters.dt)
                       - start)
        with get_progress(progress_type)(
                num_step, "--> TEMPO-with-field computation:") as prog:
            for i in range(num_step):
                prog.update(i)
                step, states, field = self._backend_instance.compute_step()
                self._dynamics.add(self._time(step), self._matrices(states),
                                   field)
            prog.update(num_step)
        return self._dynamics

    def get_dynamics(self) -> MeanFieldDynamics:
        """The mean-field dynamics computed so far."""
        return self._dynamics
