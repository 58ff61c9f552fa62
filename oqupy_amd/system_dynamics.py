"""``compute_dynamics``: system dynamics through a process tensor.

Restates ``compute_dynamics`` and its input parsing from
``oqupy/system_dynamics.py:41-182, 478-559`` for no, one or several (up to 8)
process tensors, each carrying its own bond leg of the state.
Control operations are folded into the half-step propagators on the host
(the post-measurement control of step k into the first half-step of step k,
the pre-measurement control of step k+1 into the second half-step of step k,
that of step 0 into the initial state), which reproduces the reference's
order of application. The per-step contraction -- first half-step propagator,
PT-MPO tensor (with the basis transforms of a non-diagonal coupling), second
half-step, and the cap readout of every state -- runs in the MI355X engine
(``pte_dynamics_run`` / ``pte_dynamics_run_engine`` for one process tensor,
``pte_dynamics_run_engines`` / ``pte_dynamics_run_multi`` for several).
"""

import ctypes
from typing import List, Optional, Text, Union

import numpy as np
from numpy import ndarray

from oqupy_amd.backends import engine as eng
from oqupy_amd.config import INTEGRATE_EPSREL, SUBDIV_LIMIT
from oqupy_amd.control import Control
from oqupy_amd.dynamics import Dynamics
from oqupy_amd.process_tensor import BaseProcessTensor, TrivialProcessTensor
from oqupy_amd.system import System, TimeDependentSystem
from oqupy_amd.util import check_convert, check_isinstance, check_true


def _compute_dynamics_input_parse(system, initial_state, dt, num_steps,
                                  start_time, process_tensor, control,
                                  record_all):
    """oqupy/system_dynamics.py:478-559."""
    check_isinstance(system, (System, TimeDependentSystem), "system")
    hs_dim = system.dimension
    if process_tensor is None:
        process_tensor = []
    check_isinstance(process_tensor, (BaseProcessTensor, list),
                     "process_tensor")
    process_tensors = [process_tensor] \
        if isinstance(process_tensor, BaseProcessTensor) else process_tensor
    if control is None:
        control = Control(hs_dim)
    check_isinstance(control, Control, "control")
    start_time = check_convert(start_time, float, "start_time")
    if initial_state is None:
        raise ValueError("An initial state must be specified.")
    check_isinstance(initial_state, ndarray, "initial_state")
    check_true(initial_state.shape == (hs_dim, hs_dim),
               "Initial sate must be a square matrix of "
               f"dimension {hs_dim}x{hs_dim}.")
    max_steps = []
    for pt in process_tensors:
        check_isinstance(pt, BaseProcessTensor, "pt",
                         "One of the elements in `process_tensor` is not of "
                         f"type `{BaseProcessTensor.__name__}`.")
        if pt.get_initial_tensor() is not None:
            raise NotImplementedError()
        check_true(hs_dim == pt.hilbert_space_dimension,
                   "All process tensor must have the same Hilbert space "
                   "dimension as the system.")
        if pt.dt is not None:
            if dt is None:
                dt = pt.dt
            else:
                check_true(pt.dt == dt, "All process tensors must have the "
                           "same timestep length.")
        max_steps.append(pt.max_step)
    max_step = np.min(max_steps + [np.inf])
    if dt is None:
        raise ValueError(
            "No timestep length has been specified. Please set `dt`.")
    if num_steps is not None:
        num_steps = check_convert(num_steps, int, "num_steps")
        check_true(num_steps <= max_step, "Variable `num_steps` is larger "
                   "than the shortest process tensor!")
    else:
        check_true(max_step < np.inf, "Variable `num_steps` must be "
                   "specified because all process tensors involved are "
                   "infinite.")
        num_steps = int(max_step)
    return (system, initial_state, dt, num_steps, start_time, process_tensors,
            control, record_all, hs_dim)


def _fold_controls(control: Control, props: ndarray, rho0: ndarray,
                   num_steps: int, dt: float, start_time: float) -> ndarray:
    """Fold the pre/post-measurement controls into the half-step
    propagators (in place) and the initial state (returned)."""
    pre, _ = control.get_controls(0, dt=dt, start_time=start_time)
    if pre is not None:
        rho0 = np.asarray(pre, dtype=np.complex128) @ rho0
    for step in range(num_steps):
        _, post = control.get_controls(step, dt=dt, start_time=start_time)
        pre_next, _ = control.get_controls(step + 1, dt=dt,
                                           start_time=start_time)
        if post is not None:
            props[step, 0] = props[step, 0] @ post
        if pre_next is not None:
            props[step, 1] = pre_next @ props[step, 1]
    return rho0


def compute_dynamics(
        system: Union[System, TimeDependentSystem],
        initial_state: Optional[ndarray] = None,
        dt: Optional[float] = None,
        num_steps: Optional[int] = None,
        start_time: Optional[float] = 0.0,
        process_tensor: Optional[Union[List[BaseProcessTensor],
                                       BaseProcessTensor]] = None,
        control=None,
        record_all: Optional[bool] = True,
        subdiv_limit: Optional[int] = SUBDIV_LIMIT,
        liouvillian_epsrel: Optional[float] = INTEGRATE_EPSREL,
        progress_type: Optional[Text] = None) -> Dynamics:
    """Dynamics of ``system`` coupled through ``process_tensor``
    (oqupy/system_dynamics.py:41-182). States at ``start_time + k dt`` for
    ``k = 0 .. num_steps`` (only the last one if ``record_all`` is False)."""
    (system, initial_state, dt, num_steps, start_time, process_tensors,
     control, record_all, hs_dim) = _compute_dynamics_input_parse(
         system, initial_state, dt, num_steps, start_time, process_tensor,
         control, record_all)
    process_tensors = [pt for pt in process_tensors
                       if not isinstance(pt, TrivialProcessTensor)]
    if len(process_tensors) > 8:
        raise NotImplementedError("At most 8 process tensors are supported "
                                  "by this package.")
    d2 = hs_dim ** 2
    propagators = system.get_propagators(dt, start_time, subdiv_limit,
                                         liouvillian_epsrel)
    props = np.empty((num_steps, 2, d2, d2), dtype=np.complex128)
    for step in range(num_steps):
        first, second = propagators(step)
        props[step, 0] = first
        props[step, 1] = second
    rho0 = eng.as_c128(initial_state.reshape(d2))
    rho0 = eng.as_c128(_fold_controls(control, props, rho0, num_steps, dt,
                                      start_time))
    states = np.empty((num_steps + 1, d2), dtype=np.complex128)
    eng.require_gpu()
    lib = eng.load_library()

    tins = [None if pt.transform_in is None else eng.as_c128(pt.transform_in)
            for pt in process_tensors]
    touts = [None if pt.transform_out is None
             else eng.as_c128(pt.transform_out) for pt in process_tensors]
    resident = all(getattr(pt, "engine", None) is not None
                   and num_steps == len(pt) for pt in process_tensors)
    if len(process_tensors) == 1 and resident:
        eng.check(lib.pte_dynamics_run_engine(
            process_tensors[0].engine.handle, eng.dptr(props),
            eng.dptr(tins[0]), eng.dptr(touts[0]), eng.dptr(rho0),
            eng.dptr(states)))
    elif len(process_tensors) == 1:
        tensors, bonds, caps = _host_arrays(process_tensors[0], 0, num_steps)
        _run_host(lib, d2, num_steps, bonds, tensors, caps, props, tins[0],
                  touts[0], rho0, states)
    elif process_tensors and resident:
        handles = (ctypes.c_void_p * len(process_tensors))(
            *[pt.engine.handle.value for pt in process_tensors])
        eng.check(lib.pte_dynamics_run_engines(
            handles, len(process_tensors), eng.dptr(props),
            _pointer_array(tins), _pointer_array(touts), eng.dptr(rho0),
            eng.dptr(states)))
    elif process_tensors:
        tensors, bonds, caps = [], [], []
        for index, pt in enumerate(process_tensors):
            ten, bnd, cap = _host_arrays(pt, index, num_steps)
            tensors += ten
            bonds += bnd
            caps += cap
        bonds_arr = np.ascontiguousarray(bonds, dtype=np.int32)
        eng.check(lib.pte_dynamics_run_multi(
            d2, num_steps, len(process_tensors), eng.iptr(bonds_arr),
            eng.dptr(np.ascontiguousarray(np.concatenate(tensors))),
            eng.dptr(np.ascontiguousarray(np.concatenate(caps))),
            eng.dptr(props), _pointer_array(tins), _pointer_array(touts),
            eng.dptr(rho0), eng.dptr(states)))
    else:
        ones = [np.ones(d2, dtype=np.complex128) for _ in range(num_steps)]
        caps = [np.ones(1, dtype=np.complex128)] * (num_steps + 1)
        _run_host(lib, d2, num_steps, [1] * (num_steps + 1), ones, caps,
                  props, None, None, rho0, states)

    state_list = [s.reshape(hs_dim, hs_dim) for s in states]
    if record_all:
        times = start_time + np.arange(len(state_list)) * dt
        return Dynamics(times=list(times), states=state_list)
    # the reference records one state and dates it start_time + 1 * dt
    # (oqupy/system_dynamics.py:173-180)
    return Dynamics(times=[start_time + 1 * dt], states=[state_list[-1]])


def _host_arrays(pt: BaseProcessTensor, index: int, num_steps: int):
    """Host copies of the first ``num_steps`` PT-MPO tensors ([a][b][i],
    flattened), the bond dimensions and the caps of one process tensor."""
    tensors, bonds, caps = [], [], []
    for step in range(num_steps):
        ten = pt.get_raw_mpo_tensor(step)
        if ten.ndim != 3:
            raise NotImplementedError(
                "Only rank-3 PT-MPO tensors are supported.")
        tensors.append(eng.as_c128(ten).reshape(-1))
        bonds.append(ten.shape[0])
    bonds.append(pt.get_raw_mpo_tensor(num_steps - 1).shape[1]
                 if num_steps == len(pt) else
                 pt.get_raw_mpo_tensor(num_steps).shape[0])
    for step in range(num_steps + 1):
        cap = pt.get_cap_tensor(step)
        if cap is None:
            raise ValueError(f"Process tensor {index} has no cap tensor for "
                             f"step {step}.")
        caps.append(eng.as_c128(cap).reshape(-1))
    return tensors, bonds, caps


def _pointer_array(arrays):
    """``const double**`` over complex128 arrays (NULL entries for None)."""
    ptrs = (eng._DP * len(arrays))()
    for i, arr in enumerate(arrays):
        ptrs[i] = eng.dptr(arr) if arr is not None else eng._DP()
    return ptrs


def _run_host(lib, d2, num_steps, bonds, tensors, caps, props, tin, tout,
              rho0, states) -> None:
    bonds_arr = np.ascontiguousarray(bonds, dtype=np.int32)
    ten = np.ascontiguousarray(np.concatenate(tensors))
    cap = np.ascontiguousarray(np.concatenate(caps))
    eng.check(lib.pte_dynamics_run(d2, num_steps, eng.iptr(bonds_arr),
                                   eng.dptr(ten), eng.dptr(cap),
                                   eng.dptr(props), eng.dptr(tin),
                                   eng.dptr(tout), eng.dptr(rho0),
                                   eng.dptr(states)))


def compute_dynamics_batch(
        systems: List[Union[System, TimeDependentSystem]],
        process_tensor: BaseProcessTensor,
        initial_state: Union[ndarray, List[ndarray]],
        start_time: Optional[float] = 0.0,
        record_all: Optional[bool] = True,
        subdiv_limit: Optional[int] = SUBDIV_LIMIT,
        liouvillian_epsrel: Optional[float] = INTEGRATE_EPSREL,
        control_list: Optional[List[Control]] = None
        ) -> List[Dynamics]:
    """``compute_dynamics`` of many systems through one process tensor built by
    the engine (the "many drives, one environment" use of a process tensor):
    the process tensor stays in device memory and every per-step kernel serves
    all systems. No counterpart in the reference; each entry equals
    ``compute_dynamics(system_i, initial_state_i, process_tensor=...)``
    (oqupy/system_dynamics.py:41-182) over the whole process tensor."""
    if not systems:
        return []
    backend = getattr(process_tensor, "engine", None)
    if backend is None:
        raise NotImplementedError(
            "compute_dynamics_batch needs a process tensor computed by this "
            "package's PT-TEMPO (device resident).")
    states0 = initial_state if isinstance(initial_state, (list, tuple)) \
        else [initial_state] * len(systems)
    if len(states0) != len(systems):
        raise ValueError("`initial_state` needs one entry per system.")
    if control_list is None:
        control_list = [None] * len(systems)
    if len(control_list) != len(systems):
        raise ValueError("`control_list` needs one entry per system.")
    num_steps = len(process_tensor)
    dt = process_tensor.dt
    hs_dim = process_tensor.hilbert_space_dimension
    d2 = hs_dim ** 2
    props = np.empty((len(systems), num_steps, 2, d2, d2),
                     dtype=np.complex128)
    rho0 = np.empty((len(systems), d2), dtype=np.complex128)
    for s, (system, st, ctrl) in enumerate(zip(systems, states0,
                                               control_list)):
        # the same argument checks as compute_dynamics
        ctrl = _compute_dynamics_input_parse(
            system, st, dt, num_steps, start_time, process_tensor, ctrl,
            record_all)[6]
        propagators = system.get_propagators(dt, float(start_time),
                                             subdiv_limit, liouvillian_epsrel)
        for step in range(num_steps):
            first, second = propagators(step)
            props[s, step, 0] = first
            props[s, step, 1] = second
        rho0[s] = _fold_controls(ctrl, props[s],
                                 np.asarray(st, dtype=np.complex128
                                            ).reshape(d2),
                                 num_steps, dt, float(start_time))
    tin = None if process_tensor.transform_in is None \
        else eng.as_c128(process_tensor.transform_in)
    tout = None if process_tensor.transform_out is None \
        else eng.as_c128(process_tensor.transform_out)
    states = np.empty((len(systems), num_steps + 1, d2), dtype=np.complex128)
    eng.require_gpu()
    lib = eng.load_library()
    eng.check(lib.pte_dynamics_run_engine_batch(
        backend.handle, len(systems), eng.dptr(props), eng.dptr(tin),
        eng.dptr(tout), eng.dptr(rho0), eng.dptr(states)))
    out = []
    for s in range(len(systems)):
        state_list = [x.reshape(hs_dim, hs_dim) for x in states[s]]
        if record_all:
            times = float(start_time) + np.arange(num_steps + 1) * dt
            out.append(Dynamics(times=list(times), states=state_list))
        else:
            out.append(Dynamics(times=[float(start_time) + 1 * dt],
                                states=[state_list[-1]]))
    return out
