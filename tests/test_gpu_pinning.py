"""The engine on the oracle's pinned paths (needs a GPU): the basis
rotations of a non-diagonal coupling (the reference's
tests/physics/tempo_non_diagonal_test.py:23-133) and the degeneracy-reduced
networks of unique=True (pt_tempo_backend.py:114-140,
tempo_backend.py:400-417), each against the oracle and against the
invariance itself."""

import sys

import numpy as np
import pytest

import oqupy_amd as oqupy
from oqupy_amd.tempo import TempoParameters

from tests import models
from tests.test_gpu_parity import _assert_within_spread, STATE_TOL
from tests.test_oracle_pinning import BASES, ROTATION_PARAMS, \
    ROTATION_STEPS, _rotation_model, _spin_components, oracle_pt_states, \
    oracle_tempo_states, spin_one_model

pytestmark = pytest.mark.gpu


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)
    sys.stdout.flush()


def _engine_pt_states(system, bath, params, rho0, n, unique=False):
    pt = oqupy.pt_tempo_compute(bath, 0.0, n * params.dt + 1e-9 * params.dt,
                                params, unique=unique)
    dyn = oqupy.compute_dynamics(system=system, process_tensor=pt,
                                 initial_state=rho0)
    return np.asarray(dyn.states), pt


def _engine_tempo_states(system, bath, params, rho0, n, unique=False):
    tempo = oqupy.Tempo(system, bath, params, rho0, start_time=0.0,
                        unique=unique)
    dyn = tempo.compute(end_time=n * params.dt + 1e-9 * params.dt)
    return np.asarray(dyn.states)


def test_pt_tempo_rotation_invariance_engine():
    _announce("pt_tempo rotation invariance (three bases)")
    res = []
    for i, base in enumerate(BASES):
        system, bath, params, rho0 = _rotation_model(*base)
        states, _ = _engine_pt_states(system, bath, params, rho0,
                                      ROTATION_STEPS)
        ref, _ = oracle_pt_states(system, bath, params, rho0, ROTATION_STEPS)
        assert np.abs(states - ref).max() < STATE_TOL, base
        res.append(_spin_components(states, i))
    assert np.allclose(res[0], res[1], atol=ROTATION_PARAMS["epsrel"])
    assert np.allclose(res[0], res[2], atol=ROTATION_PARAMS["epsrel"])


def test_tempo_rotation_invariance_engine():
    _announce("tempo rotation invariance (three bases)")
    from tests.test_gpu_parity import _oracle_tempo_spread
    res = []
    for i, base in enumerate(BASES):
        system, bath, params, rho0 = _rotation_model(*base)
        states = _engine_tempo_states(system, bath, params, rho0,
                                      ROTATION_STEPS)
        ref, spread = _oracle_tempo_spread(system, bath, params, rho0,
                                           ROTATION_STEPS)
        _assert_within_spread(states, ref, spread)
        res.append(_spin_components(states, i))
    assert np.allclose(res[0], res[1], atol=ROTATION_PARAMS["epsrel"])
    assert np.allclose(res[0], res[2], atol=ROTATION_PARAMS["epsrel"])


@pytest.mark.parametrize("dkmax,n,epsrel", [(3, 7, 1e-13), (4, 12, 1e-7)])
def test_pt_tempo_unique_engine(dkmax, n, epsrel):
    _announce(f"pt_tempo unique=True spin-1 dkmax={dkmax} eps={epsrel}")
    system, bath, rho0 = spin_one_model()
    params = TempoParameters(dt=0.05, dkmax=dkmax, epsrel=epsrel)
    red, pt = _engine_pt_states(system, bath, params, rho0, n, unique=True)
    ref, ref_pt = oracle_pt_states(system, bath, params, rho0, n, unique=True)
    assert list(pt.get_bond_dimensions()[1:-1]) == ref_pt.bond_dimensions()
    assert np.abs(red - ref).max() < STATE_TOL
    full, _ = _engine_pt_states(system, bath, params, rho0, n)
    tol = 1e-11 if epsrel < 1e-12 else 10 * epsrel
    assert np.abs(full - red).max() < tol


def _oracle_tempo_unique_spread(system, bath, params, rho0, n):
    """Oracle TEMPO (unique=True) states and the spread between two LAPACK
    drivers (gesdd vs gesvd) on the same network: TEMPO's truncation
    amplifies rounding-level differences (see test_gpu_parity)."""
    import scipy.linalg as sl
    from oracle import chain
    ref = oracle_tempo_states(system, bath, params, rho0, n, unique=True)
    orig = chain.truncated_svd

    def gesvd(mat, eps, spectra=None):
        u, s, vh = sl.svd(mat, full_matrices=False, lapack_driver="gesvd")
        k = chain.truncation_rank(s, eps)
        return u[:, :k], s[:k], vh[:k]
    chain.truncated_svd = gesvd
    try:
        alt = oracle_tempo_states(system, bath, params, rho0, n, unique=True)
    finally:
        chain.truncated_svd = orig
    return ref, np.abs(alt - ref).max(axis=(1, 2))


@pytest.mark.parametrize("dkmax,n,epsrel", [(3, 7, 1e-13), (4, 12, 1e-7)])
def test_tempo_unique_engine(dkmax, n, epsrel):
    _announce(f"tempo unique=True spin-1 dkmax={dkmax} eps={epsrel}")
    system, bath, rho0 = spin_one_model()
    params = TempoParameters(dt=0.05, dkmax=dkmax, epsrel=epsrel)
    red = _engine_tempo_states(system, bath, params, rho0, n, unique=True)
    ref, spread = _oracle_tempo_unique_spread(system, bath, params, rho0, n)
    print(f"  engine vs oracle {np.abs(red - ref).max():.2e}, LAPACK-driver "
          f"spread {spread.max():.2e}", flush=True)
    _assert_within_spread(red, ref, spread)
    full = _engine_tempo_states(system, bath, params, rho0, n)
    tol = 1e-11 if epsrel < 1e-12 else 10 * epsrel
    assert np.abs(full - red).max() < tol
