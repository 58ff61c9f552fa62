"""Mean-field TEMPO on explicit arrays (FP64 oracle; test infrastructure).

Restates ``MeanFieldTempoBackend`` (``oqupy/backends/tempo_backend.py:
629-773``) with the field rules of ``MeanFieldTempo``
(``oqupy/tempo.py:867-892``): every system has its own TEMPO network
(``oracle.tempo.TempoOracle``); per step the field derivative f'(t_n) =
eom(t_n, [rho_n], f_n) is evaluated with the states at the step's start, each
system advances with propagators built from (f_n, f'(t_n)), and the field
advances by Heun's rule f_{n+1} = f_n + dt/2 (k_1 + k_2) with
k_1 = eom(t_n, [rho_n], f_n), k_2 = eom(t_n + dt, [rho_{n+1}], f_n + dt k_1).

Pinning: the golden field and state of the reference's
tests/physics/mean_field_tempo_test.py:10-62 (decimal 4), and, for one
system without bath coupling strength, the closed form of the field for a
field equation that does not depend on the states.
"""

from typing import Callable, List

import numpy as np

from oracle.tempo import TempoOracle


class MeanFieldTempoOracle:
    """FP64 restatement of the mean-field TEMPO step."""

    def __init__(self, initial_states: List[np.ndarray], initial_field: complex,
                 influences: List[Callable[[int], np.ndarray]],
                 unitaries: List[np.ndarray],
                 propagators: List[Callable[[int, complex, complex], tuple]],
                 field_eom: Callable, dt: float, start_time: float,
                 dkmax, epsrel: float) -> None:
        self.dims = [rho.shape[0] for rho in initial_states]
        self.field = complex(initial_field)
        self.propagators = propagators
        self.field_eom = field_eom
        self.dt = dt
        self.start_time = start_time
        self.nets = [TempoOracle(rho, infl, u, None, np.ones(d * d),
                                 np.ones(d * d), dkmax, epsrel)
                     for rho, infl, u, d in zip(initial_states, influences,
                                                unitaries, self.dims)]
        self.states = [np.array(rho, dtype=complex) for rho in initial_states]
        self.step = None

    def initialize(self):
        for net in self.nets:
            net.initialize()
        self.step = 0
        return self.step, [s.copy() for s in self.states], self.field

    def compute_step(self):
        n = self.step
        t = self.start_time + n * self.dt
        k1 = complex(self.field_eom(t, self.states, self.field))
        new = []
        for net, prop, d in zip(self.nets, self.propagators, self.dims):
            p1, p2 = prop(n, self.field, k1)
            vec = net.compute_system_step(n + 1, p1, p2)
            new.append(vec.reshape(d, d))
        k2 = complex(self.field_eom(t + self.dt, new,
                                    self.field + self.dt * k1))
        self.field = self.field + self.dt * (k1 + k2) / 2
        self.states = new
        self.step = n + 1
        return self.step, [s.copy() for s in self.states], self.field
