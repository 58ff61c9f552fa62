"""Bath correlation integrals on the device (needs a GPU): eta(t) of
``pte_eta_batch`` against its closed form and against tight host quadrature,
and process tensors built from the device influence tensors against the
oracle fed the same tensors."""

import sys

import numpy as np
import pytest

import oqupy_amd as oqupy
from oqupy_amd import operators as op
from oqupy_amd.influence import DeviceInfluence, eta_values

from tests import models

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-10


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)
    sys.stdout.flush()


def test_eta_closed_form():
    """eta(t) = 2 alpha [ln(1 + i w_c t) - i w_c t] for an Ohmic exponential
    bath at T = 0, several baths in one launch."""
    _announce("eta closed form")
    cases = [(0.3, 3.0), (0.5, 5.0), (0.05, 10.0)]
    corrs = [oqupy.PowerLawSD(a, 1.0, wc, temperature=0.0) for a, wc in cases]
    times = [np.arange(0, 302) * 0.05, np.arange(0, 42) * 0.05,
             np.arange(0, 402) * 0.025]
    got = eta_values(corrs, times)
    for (a, wc), t, g in zip(cases, times, got):
        exact = 2 * a * (np.log(1 + 1j * wc * t) - 1j * wc * t)
        err = np.abs(g - exact).max() / np.abs(exact).max()
        print(f"  alpha {a} w_c {wc}: max error {err:.2e} (relative to max)")
        assert err < 1e-11
        assert g[0] == 0.0


@pytest.mark.parametrize("zeta,cutoff,temp", [(1.0, "exponential", 0.2),
                                              (1.0, "exponential", 0.9),
                                              (3.0, "exponential", 0.5),
                                              (1.0, "gaussian", 0.3),
                                              (2.0, "hard", 0.4)])
def test_eta_vs_host_quadrature(zeta, cutoff, temp):
    _announce(f"eta vs quadrature zeta={zeta} {cutoff} T={temp}")
    corr = oqupy.PowerLawSD(0.4, zeta, 4.0, cutoff_type=cutoff,
                            temperature=temp)
    t = np.array([0.05, 0.3, 1.0, 2.5, 6.0])
    got = eta_values([corr], [t])[0]
    ref = np.array([corr.eta_function(x, epsrel=1e-12, subdiv_limit=2000)
                    for x in t])
    err = np.abs(got - ref) / np.abs(ref)
    print(f"  max relative deviation {err.max():.2e}")
    assert err.max() < 1e-8


def test_pt_from_device_influence():
    """PT-TEMPO with backend_config {"influence": "device"} against the
    oracle fed the same device influence tensors (bond dimensions identical,
    states 1e-10), and against the host-quadrature process tensor (whose eta
    carries QUADPACK's epsrel = 1e-7 error)."""
    _announce("pt from device influence")
    system, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.05, 6, 1e-7)
    n = 20
    rho0 = op.spin_dm("z+")
    t_end = n * 0.05 + 1e-11
    pt = oqupy.pt_tempo_compute(bath, 0.0, t_end, params,
                                backend_config={"influence": "device"})
    dyn = oqupy.compute_dynamics(system=system, process_tensor=pt,
                                 initial_state=rho0)
    infl = DeviceInfluence(bath, params, 6)
    ref_states, ref_bonds, _ = models.oracle_pt_dynamics(
        system, bath, params, rho0, n, influence=infl)
    assert list(pt.get_bond_dimensions()[1:-1]) == list(ref_bonds)
    assert np.abs(dyn.states - ref_states).max() < STATE_TOL
    host = oqupy.pt_tempo_compute(bath, 0.0, t_end, params)
    dyn_h = oqupy.compute_dynamics(system=system, process_tensor=host,
                                   initial_state=rho0)
    dev = np.abs(models.sigma_z_expectation(dyn.states)
                 - models.sigma_z_expectation(dyn_h.states)).max()
    print(f"  device vs host influence: max |d<sz>| = {dev:.2e}")
    assert dev < 1e-5
    # the batched computation hands eta of all baths over in one launch
    batch = oqupy.pt_tempo_compute_batch([bath, bath], 0.0, t_end, params,
                                         backend_config={"influence":
                                                         "device"})
    dyn_b = oqupy.compute_dynamics(system=system, process_tensor=batch[1],
                                   initial_state=rho0)
    assert np.abs(dyn_b.states - dyn.states).max() < 1e-13
