"""Host-side API of the package (CPU only): input validation with the
reference's exception types (tests/coverage/tempo_test.py:76-194,
tests/coverage/pt_tempo_test.py:49-79), tensor-layout helpers and
influence tensors."""

import numpy as np
import pytest

import oqupy_amd as oqupy
from oqupy_amd import operators as op
from oqupy_amd.util import add_singleton, create_delta


def _loop_create_delta(tensor, scramble):
    out = np.zeros(tuple(tensor.shape[i] for i in scramble), dtype=tensor.dtype)
    for idx in np.ndindex(*tensor.shape):
        out[tuple(idx[i] for i in scramble)] = tensor[idx]
    return out


@pytest.mark.parametrize("scramble,shape", [([1, 1, 0], (4, 4)),
                                            ([0, 1, 1, 0], (4, 3)),
                                            ([1, 0, 0, 1], (3, 5)),
                                            ([0, 1, 0], (2, 4)),
                                            ([0, 1, 2, 2], (2, 3, 4))])
def test_create_delta_matches_definition(scramble, shape):
    rng = np.random.default_rng(1)
    t = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    assert np.array_equal(create_delta(t, scramble), _loop_create_delta(t, scramble))


def test_add_singleton():
    t = np.arange(6.0).reshape(2, 3)
    assert add_singleton(t, 1).shape == (2, 1, 3)
    assert add_singleton(add_singleton(t, 1), 3).shape == (2, 1, 3, 1)
    assert add_singleton(t, 2).shape == (2, 3, 1)


def test_tempo_parameters_validation():
    P = oqupy.TempoParameters
    with pytest.raises(AssertionError):
        P(dt=0.1, tcut=1, dkmax=10, epsrel=2e-5)
    with pytest.raises(TypeError):
        P(dt=0.1, dkmax=[], epsrel=2e-5)
    with pytest.raises(ValueError):
        P(dt=0.1, dkmax=-1, epsrel=2e-5)
    with pytest.raises(TypeError):
        P(dt=0.1, tcut='a', epsrel=2e-5)
    with pytest.raises(ValueError):
        P(dt=0.1, tcut=-1.0, epsrel=2e-5)
    with pytest.raises(TypeError):
        P(dt=None, dkmax=10, epsrel=2e-5)
    with pytest.raises(ValueError):
        P(dt=-0.1, dkmax=10, epsrel=2e-5)
    with pytest.raises(TypeError):
        P(dt=0.1, dkmax=10, epsrel='a')
    with pytest.raises(ValueError):
        P(dt=0.1, dkmax=10, epsrel=-2e-5)
    with pytest.raises(TypeError):
        P(dt=0.1, dkmax=10, epsrel=2e-5, add_correlation_time='a')
    with pytest.raises(ValueError):
        P(dt=0.1, dkmax=10, epsrel=2e-5, add_correlation_time=-1.0)
    with pytest.raises(TypeError):
        P(dt=0.1, dkmax=10, epsrel=2e-5, subdiv_limit='a')
    with pytest.raises(ValueError):
        P(dt=0.1, dkmax=10, epsrel=2e-5, subdiv_limit=-1.0)
    with pytest.raises(TypeError):
        P(dt=0.1, dkmax=10, epsrel=2e-5, liouvillian_epsrel='a')
    with pytest.raises(ValueError):
        P(dt=0.1, dkmax=10, epsrel=2e-5, liouvillian_epsrel=-1.0)
    param = P(dt=0.1, dkmax=10, epsrel=2e-5, subdiv_limit=None)
    assert np.isclose(param.tcut, param.dkmax * param.dt)
    assert "dkmax" in str(param) or "tcut" in str(param)
    assert P(dt=0.1, tcut=2.0, epsrel=1e-4).dkmax == 20
    assert P(dt=0.1, epsrel=1e-6).dkmax is None


def _custom_bath():
    corr = oqupy.CustomCorrelations(
        lambda t: (np.cos(6.0 * t) + 1j * np.sin(6.0 * t)) * np.exp(-12.0 * t))
    return oqupy.Bath(0.5 * op.sigma("z"), corr)


def test_pt_tempo_bad_input():
    bath = _custom_bath()
    param = oqupy.TempoParameters(0.1, 1.0e-5, None, 5, name="rough-A")
    ptt = oqupy.PtTempo(bath=bath, parameters=param, start_time=-0.3,
                        end_time=0.84)
    assert ptt.dimension == 2 and ptt.num_steps == 11
    with pytest.raises(AssertionError):
        oqupy.PtTempo(bath=bath, parameters=param, start_time="bla",
                      end_time=0.84)
    with pytest.raises(AssertionError):
        oqupy.PtTempo(bath=bath, parameters=param, start_time=-0.3,
                      end_time="bla")
    with pytest.raises(AssertionError):
        oqupy.PtTempo(bath=bath, parameters=param, start_time=0.0,
                      end_time=0.0001)
    with pytest.raises(ValueError):
        oqupy.PtTempo(bath=bath, parameters=param, start_time=0.0,
                      end_time=1.0, backend_config={"backend": "numpy"})


def test_tempo_bad_input():
    bath = _custom_bath()
    system = oqupy.System(0.5 * op.sigma("x"))
    param = oqupy.TempoParameters(0.1, 1.0e-5, None, 5)
    rho0 = op.spin_dm("z+")
    with pytest.raises(TypeError):
        oqupy.Tempo(system=system, bath=bath, parameters=param,
                    initial_state="bla", start_time=-0.3)
    with pytest.raises(TypeError):
        oqupy.Tempo(system=system, bath=bath, parameters=param,
                    initial_state=rho0, start_time="bla")
    with pytest.raises(AssertionError):
        oqupy.Tempo(system=system, bath=bath, parameters=param,
                    initial_state=rho0, start_time=-0.3, unique="bla")
    tempo = oqupy.Tempo(system=system, bath=bath, parameters=param,
                        initial_state=rho0, start_time=-0.3)
    with pytest.raises(TypeError):
        tempo.compute(end_time="bla")


def test_bath_diagonalisation_and_maps():
    corr = oqupy.PowerLawSD(0.1, 1.0, 1.0)
    bath = oqupy.Bath(0.5 * op.sigma("x"), corr)
    u = bath.unitary_transform
    assert np.allclose(u @ bath.coupling_operator @ u.conj().T,
                       0.5 * op.sigma("x"))
    bath = oqupy.Bath(0.5 * op.sigma("z"), corr)
    assert np.allclose(bath.coupling_comm, [0, 1, -1, 0])
    assert np.allclose(bath.coupling_acomm, [1, 0, 0, -1])
    assert len(set(bath.north_degeneracy_map)) == 4
    assert len(set(bath.west_degeneracy_map)) == 3
    with pytest.raises(AssertionError):
        oqupy.Bath(np.array([[0, 1], [0, 0]]), corr)


def test_influence_matrix_forms():
    corr = oqupy.PowerLawSD(0.3, 1.0, 3.0, temperature=0.0)
    bath = oqupy.Bath(0.5 * op.sigma("z"), corr)
    params = oqupy.TempoParameters(dt=0.05, dkmax=4, epsrel=1e-7)
    m, p = bath.coupling_comm, bath.coupling_acomm
    i0 = oqupy.influence_matrix(0, params, corr, p, m)
    eta0 = corr.correlation_2d_integral(0.05, 0.0, shape="upper-triangle",
                                        epsrel=1e-7)
    assert np.allclose(i0, np.diag(np.exp(-m * (eta0.real * m
                                                + 1j * eta0.imag * p))))
    i2 = oqupy.influence_matrix(2, params, corr, p, m)
    assert i2.shape == (4, 4)
    # a diagonal path index (m = 0) carries no influence
    assert np.allclose(i2[:, 0], 1.0) and np.allclose(i2[:, 3], 1.0)
    assert oqupy.influence_matrix(-3, params, corr, p, m) is None


def test_dynamics_container():
    dyn = oqupy.Dynamics(times=[0.2, 0.1], states=[op.spin_dm("z+"),
                                                    op.spin_dm("z-")])
    t, sz = dyn.expectations(op.sigma("z"), real=True)
    assert np.allclose(t, [0.1, 0.2]) and np.allclose(sz, [-1.0, 1.0])
    with pytest.raises(AssertionError):
        dyn.add(0.3, np.ones((3, 3)))


def test_default_config_does_not_cap_the_bond():
    """The reference never caps a bond dimension; the engine's bond limit is
    opt-in (chi_cap), the default is the largest C int."""
    from oqupy_amd.backends.engine import parse_backend_config
    from oqupy_amd.config import PT_TEMPO_BACKEND_CONFIG, TEMPO_BACKEND_CONFIG
    for cfg in (None, {}, PT_TEMPO_BACKEND_CONFIG, TEMPO_BACKEND_CONFIG):
        assert parse_backend_config(cfg)["chi_cap"] == 2**31 - 1
    assert parse_backend_config({"chi_cap": None})["chi_cap"] == 2**31 - 1
    assert parse_backend_config({"chi_cap": 64})["chi_cap"] == 64
    with pytest.raises(ValueError):
        parse_backend_config({"chi_cap": 0})
