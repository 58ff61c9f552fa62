"""Pinning of the FP64 oracle (CPU only).

Each oracle function is checked against something that does not restate it:
the golden states of the reference's physics tests, the exact
independent-boson result, the brute-force path sum, the closed-form eta of an
Ohmic bath at T=0, LAPACK, and invariants (trace, TEMPO == PT-TEMPO, two
baths at alpha == one at 2 alpha).
"""

import numpy as np
import pytest

from oqupy_amd import operators as op
from oqupy_amd.bath_correlations import PowerLawSD
from oqupy_amd.tempo import TempoParameters
from oracle import chain
from oracle.dynamics import compute_dynamics
from oracle.path_sum import path_sum
from oracle.pt_tempo import PtTempoOracle, compute_caps

from tests import models


@pytest.mark.parametrize("name", ["rho_A", "rho_C", "rho_D", "rho_E"])
def test_oracle_pt_tempo_golden(name):
    system, bath, params, rho0, n, rho, dec = models.golden_model(name)
    states, _, _ = models.oracle_pt_dynamics(system, bath, params, rho0, n)
    np.testing.assert_almost_equal(states[-1], rho, decimal=dec)


@pytest.mark.parametrize("name", ["rho_A", "rho_C", "rho_D", "rho_E"])
def test_oracle_tempo_golden(name):
    system, bath, params, rho0, n, rho, dec = models.golden_model(name)
    states, _ = models.oracle_tempo(system, bath, params, rho0, n)
    np.testing.assert_almost_equal(states[-1], rho, decimal=dec)


@pytest.mark.parametrize("dkmax,tau", [(3, None), (2, 0.1), (6, None),
                                       (1, 0.3)])
def test_oracles_match_path_sum(dkmax, tau):
    """At epsrel -> 0 both chain oracles equal the untruncated path sum, in
    the growth phase, the end phase and with the rectangle influences."""
    system, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.1, dkmax, 1e-14,
                                             add_correlation_time=tau)
    n = 6
    rho0 = op.spin_dm("z+")
    infl = models.CachedInfluence(bath, params)
    p1, p2 = system.get_propagators(0.1, 0.0, None, None)(0)
    exact = path_sum(rho0, p1, p2, infl, n, dkmax)
    tempo, _ = models.oracle_tempo(system, bath, params, rho0, n, infl)
    pt, _, _ = models.oracle_pt_dynamics(system, bath, params, rho0, n, infl)
    assert np.abs(tempo[-1] - exact).max() < 1e-12
    assert np.abs(pt[-1] - exact).max() < 1e-12
    assert abs(np.trace(exact) - 1) < 1e-12


def test_path_sum_detects_a_dropped_influence():
    """A plausible mistake (the dk = dkmax influence left out) changes the
    path sum far beyond the tolerance used above."""
    system, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.1, 3, 1e-14)
    infl = models.CachedInfluence(bath, params)
    rho0 = op.spin_dm("z+")
    p1, p2 = system.get_propagators(0.1, 0.0, None, None)(0)
    exact = path_sum(rho0, p1, p2, infl, 6, 3)
    dropped = path_sum(rho0, p1, p2, infl, 6, 2)
    assert np.abs(exact - dropped).max() > 1e-6


def test_truncation_rule_matches_reference_restatement():
    """tensornetwork's relative rule equals the in-tree restatement of the
    reference (oqupy/mps_mpo.py:452-457) on spectra built around the cut."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = rng.integers(1, 40)
        s = np.sort(10.0 ** rng.uniform(-12, 0, n))[::-1]
        eps = 10.0 ** rng.uniform(-9, -1)
        absrel = s[0] * eps
        ref = np.count_nonzero(np.cumsum(np.flip(s) ** 2) > absrel ** 2)
        assert chain.truncation_rank(s, eps) == ref
    # exact boundary: tail equal to the threshold is discarded
    s = np.array([1.0, 0.6, 0.8e-3])
    assert chain.truncation_rank(s, 0.8e-3) == 2
    assert chain.truncation_rank(s, 0.79e-3) == 3


def test_eta_closed_form_ohmic_zero_temperature():
    """eta(t) = 2 alpha [ln(1 + i wc t) - i wc t] for J = 2 alpha w e^{-w/wc}
    at T = 0 (from oqupy/bath_correlations.py:560-562)."""
    corr = PowerLawSD(alpha=0.3, zeta=1.0, cutoff=3.0,
                      cutoff_type="exponential", temperature=0.0)
    for t in (0.05, 0.4, 2.0, 7.5):
        exact = 2 * 0.3 * (np.log(1 + 1j * 3.0 * t) - 1j * 3.0 * t)
        got = corr.eta_function(t, epsrel=1e-12)
        assert abs(got - exact) < 1e-9 * max(1.0, abs(exact))


def test_dephasing_exact_result_ohmic():
    """With H_S = 0 and S = sigma_z/2 the coherence decays as
    exp(-Re eta(t)) (independent-boson model); PT-TEMPO + dynamics must
    reproduce it for an Ohmic bath at T = 0."""
    corr = PowerLawSD(alpha=0.2, zeta=1.0, cutoff=4.0,
                      cutoff_type="exponential", temperature=0.0)
    from oqupy_amd.bath import Bath
    from oqupy_amd.system import System
    bath = Bath(0.5 * op.sigma("z"), corr)
    system = System(0.0 * op.sigma("x"))
    n, dt = 12, 0.1
    params = TempoParameters(dt=dt, dkmax=n, epsrel=1e-10)
    rho0 = op.spin_dm("x+")
    states, _, _ = models.oracle_pt_dynamics(system, bath, params, rho0, n)
    for k in (4, 8, 12):
        t = k * dt
        exact = 0.5 * np.exp(-0.2 * np.log(1 + (4.0 * t) ** 2))
        assert abs(states[k][0, 1] - exact) < 2e-6


def test_tempo_equals_pt_tempo_and_trace():
    system, bath, params = models.spin_boson(0.3, 3.0, 0.0, 0.05, 8, 1e-7)
    rho0 = op.spin_dm("z+")
    infl = models.CachedInfluence(bath, params)
    a, _ = models.oracle_tempo(system, bath, params, rho0, 24, infl)
    b, _, _ = models.oracle_pt_dynamics(system, bath, params, rho0, 24, infl)
    assert np.abs(a - b).max() < 1e-5
    # truncation at epsrel breaks trace preservation only at that order
    assert np.abs(np.trace(b, axis1=1, axis2=2) - 1).max() < 1e-4


def test_two_baths_equal_one_at_double_coupling():
    """Two process tensors at alpha equal one at 2 alpha
    (tests/physics/multi_environments_test.py:61-68)."""
    g = models.goldens()["multi_env"]
    from oqupy_amd.bath import Bath
    from oqupy_amd.system import System
    out = {}
    for key in ("alpha", "alpha2"):
        corr = PowerLawSD(alpha=g[key], zeta=1.0, cutoff=g["cutoff"],
                          cutoff_type="exponential",
                          temperature=g["temperature"])
        bath = Bath(0.5 * op.sigma("z"), corr)
        params = TempoParameters(dt=g["dt"], dkmax=g["dkmax"],
                                 epsrel=g["epsrel"])
        infl = models.CachedInfluence(bath, params)
        pt = PtTempoOracle(2, infl, np.ones(4), g["num_steps"], g["dkmax"],
                           g["epsrel"])
        pt.compute()
        tens = pt.mpo_tensors()
        out[key] = (tens, compute_caps(tens, 2))
    system = System(op.sigma("x"))
    props = system.get_propagators(g["dt"], 0.0, None, None)
    rho0 = op.spin_dm("z+")
    two = compute_dynamics(props, rho0, g["num_steps"],
                           [out["alpha"][0], out["alpha"][0]],
                           [out["alpha"][1], out["alpha"][1]])
    one = compute_dynamics(props, rho0, g["num_steps"], [out["alpha2"][0]],
                           [out["alpha2"][1]])
    np.testing.assert_almost_equal(np.array(two), np.array(one),
                                   decimal=g["decimal"])


def test_pt_length_and_steps():
    """len(pt) == int((end - start) / dt) (oqupy/pt_tempo.py:148;
    tests/coverage/pt_tempo_test.py:40)."""
    g = models.goldens()["pt_len"]
    n = int((g["end_time"] - g["start_time"]) / g["dt"])
    assert n == g["length"]
    system, bath, params = models.spin_boson(0.1, 3.0, 0.0, g["dt"], 5, 1e-5)
    pt = PtTempoOracle(2, models.CachedInfluence(bath, params), np.ones(4),
                       n, 5, 1e-5)
    pt.compute()
    assert len(pt.mpo_tensors()) == g["length"]


def test_zip_and_sweep_match_dense_contraction():
    """With no truncation the zip-up reproduces the dense MPO-MPS product and
    the sweeps leave the state unchanged (checked on random tensors)."""
    rng = np.random.default_rng(0)

    def rand(*shape):
        return rng.normal(size=shape) + 1j * rng.normal(size=shape)

    mps = [rand(1, 2, 3), rand(3, 2, 4), rand(4, 2, 1)]
    mpo = [rand(1, 2, 3, 2), rand(2, 2, 2, 1)]
    full_before = np.einsum('apb,bqc,crd->pqr', *mps)
    chain.zip_up_left(mps, mpo, 1e-15)
    full_after = np.einsum('apb,bqc,crd->pqr', *mps)
    ref = np.einsum('pqr,qsv,vrt->pst', full_before, mpo[0][0],
                    mpo[1][..., 0])
    assert np.allclose(full_after, ref)
    chain.sweep_left(mps, 2, 0, 1e-15)
    chain.sweep_right(mps, 0, 2, 1e-15)
    assert np.allclose(np.einsum('apb,bqc,crd->pqr', *mps), ref)


# ---- controls (oqupy/control.py, system_dynamics.py:131-154) ----

def test_control_bookkeeping():
    from oqupy_amd.control import Control
    c = Control(2)
    a, b = 2.0 * np.eye(4), 3.0 * np.diag([1.0, 2.0, 3.0, 4.0])
    c.add_single(3, a)
    c.add_single(3, b)                    # composed: b @ a
    c.add_single(0.1, 5.0 * np.eye(4), post=True)   # time 0.1 -> step 2
    pre, post = c.get_controls(3, dt=0.05)
    assert np.allclose(pre, b @ a) and post is None
    pre, post = c.get_controls(2, dt=0.05)
    assert pre is None and np.allclose(post, 5.0 * np.eye(4))
    assert c.get_controls(2, dt=0.05, start_time=0.05) == (None, None)
    with pytest.raises(TypeError):
        c.add_single("3", a)
    with pytest.raises(NotImplementedError):
        c.add_continuous(lambda rho, t: rho)


@pytest.mark.parametrize("post", [False, True])
def test_oracle_controls_closed_form(post):
    """A pi pulse about x on a precessing spin (no environment): the oracle's
    placement of pre/post-measurement controls against exact unitary
    evolution."""
    import scipy.linalg as sl
    from oracle.dynamics import compute_dynamics
    from oqupy_amd import operators as op
    from oqupy_amd.control import Control
    from oqupy_amd.system import System
    dt, n, k = 0.1, 8, 3
    h = 0.7 * op.sigma("z") + 0.2 * op.sigma("x")
    system = System(h)
    props = system.get_propagators(dt, 0.0, None, None)
    x = op.sigma("x")
    pulse = op.left_right_super(x, x.conj().T)
    control = Control(2)
    control.add_single(k, pulse, post=post)
    rho0 = op.spin_dm("x+")
    states = compute_dynamics(props, rho0, n, [], [],
                              controls=lambda s: control.get_controls(s, dt=dt))
    u = lambda t: sl.expm(-1j * h * t)
    for j in range(n + 1):
        if j < k or (post and j == k):
            ref = u(j * dt) @ rho0 @ u(j * dt).conj().T
        else:
            mid = x @ u(k * dt) @ rho0 @ u(k * dt).conj().T @ x
            ref = u((j - k) * dt) @ mid @ u((j - k) * dt).conj().T
        assert np.abs(states[j] - ref).max() < 1e-12, (j, post)


@pytest.mark.parametrize("n,dkmax,split", [(12, 6, 4), (9, 5, 7)])
def test_load_state_continues_a_run(n, dkmax, split):
    """The oracle continued from another run's MPS after `split` steps (the
    benchmark's CPU leg continues from the engine's state) equals the run
    itself, in the growth phase and across the start of the end phase."""
    from oracle.pt_tempo import PtTempoOracle
    _, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.05, dkmax, 1e-7)
    infl = models.CachedInfluence(bath, params)
    whole = PtTempoOracle(2, infl, np.ones(4), n, dkmax, 1e-7)
    whole.compute()
    part = PtTempoOracle(2, infl, np.ones(4), n, dkmax, 1e-7)
    part.initialize()
    for _ in range(split - 1):
        part.compute_step()
    cont = PtTempoOracle(2, infl, np.ones(4), n, dkmax, 1e-7)
    cont.load_state(split, [s.copy() for s in part.mps])
    while cont.compute_step():
        pass
    assert cont.bond_dimensions() == whole.bond_dimensions()
    for a, b in zip(cont.mpo_tensors(), whole.mpo_tensors()):
        assert np.array_equal(a, b)


def test_oracle_spectra_log_matches_bond_log():
    """The spectra log records every truncated SVD of a step in the order of
    the bond log, with the kept counts of the truncation rule."""
    from oracle.chain import truncation_rank
    from oracle.pt_tempo import PtTempoOracle
    _, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.05, 5, 1e-7)
    infl = models.CachedInfluence(bath, params)
    pt = PtTempoOracle(2, infl, np.ones(4), 10, 5, 1e-7)
    pt.initialize()
    pt.spectra = []
    pt.compute_step()
    pt.compute_step()
    kept = [truncation_rank(s, 1e-7) for _, _, s in pt.spectra]
    assert kept == pt.bond_log[-2] + pt.bond_log[-1]
    assert all(np.all(np.diff(s) <= 0) for _, _, s in pt.spectra)
