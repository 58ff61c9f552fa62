"""Pinning of the oracle's remaining paths (CPU): basis rotations of a
non-diagonal coupling and the degeneracy-reduced (unique=True) networks.

* Rotation invariance (the reference's tests/physics/tempo_non_diagonal_test.py
  :23-133): the same spin problem written in three bases (system operator,
  coupling and initial state rotated together) gives the same rotated
  expectation values within epsrel, for TEMPO (the coupling's eigenbasis
  enters through the MPO's basis change) and for PT-TEMPO + compute_dynamics
  (through the process tensor's transforms, built by the oracle itself from
  Bath.unitary_transform). A wrong convention of either (a missing
  transpose, in and out swapped, U for U^H) breaks the invariance; the
  planted-slip test below shows it.
* unique=True vs unique=False (the reference's
  tests/physics/degeneracy_mean_field_test.py:233-256): a spin-1 with
  coupling S_z, whose Liouville indices fall into degenerate classes; the
  reduced network gives the same states as the full one.
"""

import numpy as np
import pytest

from oqupy_amd import operators as op
from oqupy_amd.bath import Bath
from oqupy_amd.bath_correlations import PowerLawSD
from oqupy_amd.system import System
from oqupy_amd.tempo import TempoParameters, degeneracy_positions, \
    influence_matrix
from oracle.dynamics import compute_dynamics, pt_transforms
from oracle.pt_tempo import PtTempoOracle, compute_caps
from oracle.tempo import TempoOracle
from tests import models

BASES = [("x", "z", "y+"), ("y", "x", "z+"), ("z", "y", "x+")]
ROTATION_PARAMS = dict(dt=0.1, dkmax=30, epsrel=1e-5, add_correlation_time=8.0)
ROTATION_STEPS = 10          # end time 1.0


def _rotation_model(sys_axis, coupling_axis, init):
    corr = PowerLawSD(alpha=0.3, zeta=1, cutoff=5.0, cutoff_type="exponential")
    bath = Bath(0.5 * op.sigma(coupling_axis), corr)
    system = System(0.5 * op.sigma(sys_axis))
    params = TempoParameters(**ROTATION_PARAMS)
    return system, bath, params, op.spin_dm(init)


def _spin_components(states, index):
    """(s_x, s_y, s_z) of every state, listed in the order of basis
    `index` (the reference's results arrays)."""
    s = [np.real(np.einsum("ij,tji->t", 0.5 * op.sigma(a), states))
         for a in "xyz"]
    order = [(0, 1, 2), (1, 2, 0), (2, 0, 1)][index]
    return np.array([s[i] for i in order])


def oracle_tempo_states(system, bath, params, rho0, n, unique=False):
    infl = models.CachedInfluence(bath, params)
    maps, d = None, rho0.shape[0]
    sum_n = sum_w = np.ones(d * d)
    if unique:
        pos = degeneracy_positions(bath)

        def infl(dk, _p=params):
            return influence_matrix(dk, _p, bath.correlations,
                                    bath.coupling_acomm, bath.coupling_comm,
                                    deg_positions=pos)
        maps = [bath.north_degeneracy_map, bath.west_degeneracy_map]
        sum_n = np.ones(np.max(maps[0]) + 1)
        sum_w = np.ones(np.max(maps[1]) + 1)
    props = system.get_propagators(params.dt, 0.0, None, None)
    tem = TempoOracle(rho0, infl, bath.unitary_transform, props, sum_n, sum_w,
                      params.dkmax, params.epsrel, maps, d)
    _, st = tem.initialize()
    out = [st.reshape(d, d)]
    for _ in range(n):
        _, st = tem.compute_step()
        out.append(st.reshape(d, d))
    return np.array(out)


def oracle_pt_states(system, bath, params, rho0, n, unique=False,
                     transforms=None):
    d = rho0.shape[0]
    infl = models.CachedInfluence(bath, params)
    maps, sum_n = None, np.ones(d * d)
    if unique:
        pos = degeneracy_positions(bath)

        def infl(dk, _p=params):
            return influence_matrix(dk, _p, bath.correlations,
                                    bath.coupling_acomm, bath.coupling_comm,
                                    deg_positions=pos)
        maps = [bath.north_degeneracy_map, bath.west_degeneracy_map]
        sum_n = np.ones(np.max(maps[0]) + 1)
    pt = PtTempoOracle(d, infl, sum_n, n, params.dkmax, params.epsrel, maps)
    pt.compute()
    tens = pt.mpo_tensors()
    caps = compute_caps(tens, d)
    if transforms is None:
        transforms = pt_transforms(bath.unitary_transform)
    props = system.get_propagators(params.dt, 0.0, None, None)
    return np.array(compute_dynamics(props, rho0, n, [tens], [caps],
                                     transforms=[transforms])), pt


def test_tempo_rotation_invariance():
    res = []
    for i, base in enumerate(BASES):
        system, bath, params, rho0 = _rotation_model(*base)
        states = oracle_tempo_states(system, bath, params, rho0,
                                     ROTATION_STEPS)
        res.append(_spin_components(states, i))
    assert np.allclose(res[0], res[1], atol=ROTATION_PARAMS["epsrel"])
    assert np.allclose(res[0], res[2], atol=ROTATION_PARAMS["epsrel"])
    # the coupling is non-diagonal in two of the bases: the dynamics is not
    # trivially the same
    assert np.ptp(res[0]) > 0.1


def test_pt_tempo_rotation_invariance():
    res = []
    for i, base in enumerate(BASES):
        system, bath, params, rho0 = _rotation_model(*base)
        states, _ = oracle_pt_states(system, bath, params, rho0,
                                     ROTATION_STEPS)
        res.append(_spin_components(states, i))
    assert np.allclose(res[0], res[1], atol=ROTATION_PARAMS["epsrel"])
    assert np.allclose(res[0], res[2], atol=ROTATION_PARAMS["epsrel"])


@pytest.mark.parametrize("slip", ["swap", "no_transpose"])
def test_rotation_invariance_catches_a_wrong_transform(slip):
    """A planted mistake in the transforms (in and out swapped, or the
    transpose of the leg contraction dropped) breaks the invariance."""
    res = []
    for i, base in enumerate(BASES):
        system, bath, params, rho0 = _rotation_model(*base)
        t_in, t_out = pt_transforms(bath.unitary_transform)
        if t_in is not None:
            t_in, t_out = (t_out, t_in) if slip == "swap" else (t_in.T,
                                                                  t_out.T)
        states, _ = oracle_pt_states(system, bath, params, rho0,
                                     ROTATION_STEPS, transforms=(t_in, t_out))
        res.append(_spin_components(states, i))
    assert not (np.allclose(res[0], res[1], atol=1e-3)
                and np.allclose(res[0], res[2], atol=1e-3))


def spin_one_model():
    """Spin 1 (the reference's degeneracy test, case E): H = (J/2) S_z^2 +
    h S_x, coupling S_z (eigenvalues 1, 0, -1: 5 distinct differences over
    the 9 Liouville indices, so 5 west classes)."""
    s_z = np.diag([1.0, 0.0, -1.0])
    s_x = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 1.0], [0.0, 1.0, 0.0]])
    system = System(-0.5 * s_z @ s_z + 1.0 * s_x)
    corr = PowerLawSD(alpha=0.3, zeta=1.0, cutoff=5.0,
                      cutoff_type="exponential", temperature=0.2)
    bath = Bath(s_z, corr)
    rho0 = np.diag([1.0, 0.0, 0.0]).astype(complex)
    return system, bath, rho0


@pytest.mark.parametrize("form", ["tempo", "pt_tempo"])
@pytest.mark.parametrize("dkmax,n,epsrel,tol", [(3, 7, 1e-13, 1e-11),
                                                (4, 12, 1e-7, 1e-6)])
def test_unique_matches_full_network(form, dkmax, n, epsrel, tol):
    """Without truncation (epsrel 1e-13) the reduced network is an exact
    rewriting of the full one: states agree to rounding. With truncation the
    two networks are truncated differently (the reduced SVD matrices weigh
    each class once), so they agree to the truncation error, 10 epsrel."""
    system, bath, rho0 = spin_one_model()
    assert np.max(bath.west_degeneracy_map) + 1 == 5
    params = TempoParameters(dt=0.05, dkmax=dkmax, epsrel=epsrel)
    if form == "tempo":
        full = oracle_tempo_states(system, bath, params, rho0, n)
        red = oracle_tempo_states(system, bath, params, rho0, n, unique=True)
    else:
        full, _ = oracle_pt_states(system, bath, params, rho0, n)
        red, _ = oracle_pt_states(system, bath, params, rho0, n, unique=True)
    err = np.abs(full - red).max()
    assert err < tol, err


def test_rounding_ambiguous_cut_rule():
    """The parity rule for a kept rank that differs from the oracle's: only
    a rank the cut reaches with the spectrum moved by rel_tol s_0."""
    from oracle.chain import truncation_rank
    eps, tol = 1e-3, 1e-10
    # tail norm from index 2 on is exactly at the threshold (up to rounding)
    s = np.array([1.0, 0.5, eps * 0.8, eps * 0.6])
    assert truncation_rank(s, eps) in (2, 3)
    assert [k2 for k2 in range(5)
            if models.admissible_rank(s, k2, eps, tol)] == [2, 3]
    # a clear gap: only the rule's own rank
    s = np.array([1.0, 0.5, 1e-2, 1e-6, 1e-7])
    k = truncation_rank(s, eps)
    assert k == 3
    assert [k2 for k2 in range(6)
            if models.admissible_rank(s, k2, eps, tol)] == [3]


def test_steered_cut_adopts_only_ambiguous_ranks():
    """SteeredCut on an oracle run: the oracle's own log changes nothing; a
    log whose rank is off by one at a cut with a clear gap is not adopted,
    so the comparison fails."""
    system, bath, params = models.spin_boson(0.3, 3.0, 0.0, 0.05, 4, 1e-4)
    infl = models.CachedInfluence(bath, params)
    ref = PtTempoOracle(2, infl, np.ones(4), 6, 4, 1e-4)
    ref.spectra = []
    ref.compute()
    from oracle.chain import truncation_rank
    own = [(m, n, s[:truncation_rank(s, 1e-4)]) for m, n, s in ref.spectra]
    again = PtTempoOracle(2, infl, np.ones(4), 6, 4, 1e-4)
    again.spectra = []
    with models.SteeredCut(own, 1e-4, 1e-10) as steer:
        again.compute()
    assert steer.adopted == []
    models.compare_svd_logs(own, again.spectra, 1e-4)
    # plant a rank one short at the first SVD that keeps more than one value
    i = next(j for j, (_, _, kept) in enumerate(own) if len(kept) > 1)
    slipped = list(own)
    slipped[i] = (own[i][0], own[i][1], own[i][2][:-1])
    third = PtTempoOracle(2, infl, np.ones(4), 6, 4, 1e-4)
    third.spectra = []
    with models.SteeredCut(slipped, 1e-4, 1e-10) as steer:
        third.compute()
    assert i not in steer.adopted
    with pytest.raises(AssertionError):
        models.compare_svd_logs(slipped, third.spectra, 1e-4,
                                adopted=steer.adopted)
