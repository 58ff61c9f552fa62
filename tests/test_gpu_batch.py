"""Batched PT-TEMPO (needs a GPU): many instances with different baths,
memory cuts and lengths stepped in lockstep through batched SVD launches
must give exactly what the FP64 oracle gives for each instance alone:
identical bond dimensions, states within 1e-10."""

import sys

import numpy as np
import pytest

import oqupy_amd as oqupy
from oqupy_amd import operators as op

from tests import models

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-10

# (alpha, cutoff, temperature, dkmax, num_steps)
CASES = [(0.5, 5.0, 0.2, 6, 16), (0.1, 3.0, 0.0, 4, 12), (0.8, 2.0, 1.0, 8, 20),
         (0.3, 8.0, 0.5, 3, 9), (0.05, 1.5, 0.1, 10, 14), (0.6, 5.0, 0.0, 5, 18)]


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)
    sys.stdout.flush()


def test_batch_matches_oracle_per_instance():
    _announce(f"batch of {len(CASES)} instances")
    dt, epsrel = 0.05, 1e-7
    models_ = [models.spin_boson(a, wc, t, dt, k, epsrel)
               for a, wc, t, k, _ in CASES]
    baths = [m[1] for m in models_]
    params = [m[2] for m in models_]
    ends = [n * dt + 1e-11 for *_, n in CASES]
    pts = oqupy.pt_tempo_compute_batch(baths, 0.0, ends, params)
    rho0 = op.spin_dm("z+")
    for (system, bath, prm), pt, case in zip(models_, pts, CASES):
        n = case[-1]
        assert len(pt) == n
        dyn = oqupy.compute_dynamics(system=system, process_tensor=pt,
                                     initial_state=rho0)
        ref_states, ref_bonds, _ = models.oracle_pt_dynamics(system, bath, prm,
                                                             rho0, n)
        assert list(pt.get_bond_dimensions()[1:-1]) == ref_bonds, case
        assert np.abs(dyn.states - ref_states).max() < STATE_TOL, case


def test_batch_of_one_equals_single():
    _announce("batch of one")
    system, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.05, 6, 1e-7)
    end = 14 * 0.05 + 1e-11
    (pt_b,) = oqupy.pt_tempo_compute_batch([bath], 0.0, end, params)
    pt_s = oqupy.pt_tempo_compute(bath, 0.0, end, params)
    # the batched Jacobi and the single engine's LDS-tile Jacobi round
    # differently, so the factors agree up to the gauge of the SVDs: compare
    # the bonds and the dynamics the two process tensors give
    assert list(pt_b.get_bond_dimensions()) == list(pt_s.get_bond_dimensions())
    rho0 = op.spin_dm("z+")
    dyn_b = oqupy.compute_dynamics(system=system, process_tensor=pt_b,
                                   initial_state=rho0)
    dyn_s = oqupy.compute_dynamics(system=system, process_tensor=pt_s,
                                   initial_state=rho0)
    assert np.abs(np.asarray(dyn_b.states) - np.asarray(dyn_s.states)).max() < 1e-12


def test_batch_rejects_unsupported():
    system, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.05, 4, 1e-7,
                                             add_correlation_time=0.1)
    with pytest.raises(NotImplementedError):
        oqupy.pt_tempo_compute_batch([bath], 0.0, 0.5, params)
    _, bath2, params2 = models.spin_boson(0.5, 5.0, 0.2, 0.05, 4, 1e-7)
    with pytest.raises(ValueError):
        oqupy.pt_tempo_compute_batch([bath2, bath2], 0.0, [0.5], params2)


def test_dynamics_batch_many_systems_one_pt():
    """compute_dynamics_batch equals compute_dynamics per system."""
    _announce("dynamics batch: 5 systems, one process tensor")
    _, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.05, 6, 1e-7)
    n = 16
    pt = oqupy.pt_tempo_compute(bath, 0.0, n * 0.05 + 1e-11, params)
    systems = [oqupy.System(0.5 * om * op.sigma("x")) for om in (0.5, 1.0, 2.0)]
    systems.append(oqupy.System(0.5 * op.sigma("x") + 0.3 * op.sigma("z")))
    systems.append(oqupy.TimeDependentSystem(
        lambda t: 0.5 * np.cos(2.0 * t) * op.sigma("x")))
    rhos = [op.spin_dm("z+"), op.spin_dm("x+"), op.spin_dm("z-"),
            op.spin_dm("y+"), op.spin_dm("z+")]
    batch = oqupy.compute_dynamics_batch(systems, pt, rhos)
    for system, rho0, dyn_b in zip(systems, rhos, batch):
        dyn = oqupy.compute_dynamics(system=system, process_tensor=pt,
                                     initial_state=rho0)
        assert np.abs(np.asarray(dyn_b.states) - np.asarray(dyn.states)).max() \
            < 1e-13
        assert np.allclose(dyn_b.times, dyn.times)
