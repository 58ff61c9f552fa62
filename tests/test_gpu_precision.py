"""The precision form (needs a GPU): network contractions in complex64 on the
matrix cores with FP32 accumulation, QR / Jacobi SVD / singular values / cut
in FP64. Its results are compared with the FP64 oracle and the deviation is
printed; the bounds below are the stated accuracy of the form, not parity
(the plain complex128 form is held to 1e-10 in test_gpu_parity.py).
"""

import sys

import numpy as np
import pytest

import oqupy_amd as oqupy
from oqupy_amd import operators as op
from oqupy_amd.backends import engine

from tests import models

pytestmark = pytest.mark.gpu

# single-precision contractions: |C - AB| <= K * 2^-24 * |A| |B| per entry;
# the bound leaves a factor ~10 over that worst case
C64_GEMM_REL = 1e-5
# max |<sigma_z>_c64 - <sigma_z>_oracle| over the trajectory: the form's
# stated accuracy
C64_SIGMA_Z_TOL = 1e-6
# max |tr rho - 1|: the complex64 rounding of every contraction operand
# (unit roundoff 6e-8) scales the state by 1 + O(1e-7) at every site of
# every step; measured 3.1e-6 on the 24-step case below, the bound leaves a
# factor ~3
C64_TRACE_TOL = 1e-5
# TEMPO carries the physical state itself through a complex64 contraction at
# every site of every step (PT-TEMPO only its process tensor, which
# compute_dynamics then contracts in FP64): the rounding accumulates to a
# measured 1.22e-6 on <sigma_z> of the rho_A case (20 steps, no memory
# cut, epsrel 1e-7), against a spread of 5.7e-8 between two FP64 LAPACK
# drivers; the form's TEMPO floor
C64_TEMPO_SIGMA_Z_TOL = 2e-6


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 3, 7), (70, 90, 45),
                                   (130, 64, 200), (64, 129, 16)])
def test_contract_kernels(m, n, k):
    _announce(f"contract {m}x{n}x{k}")
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    a = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    b = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    ref = a @ b
    scale = np.abs(a).max() * np.abs(b).max() * k
    c128 = engine.contract(a, b, "c128")
    assert np.abs(c128 - ref).max() <= 1e-13 * scale
    c64 = engine.contract(a, b, "c64")
    err = np.abs(c64 - ref).max() / scale
    print(f"  c64 contraction error {err:.2e} (relative to K max|A| max|B|)")
    assert err <= C64_GEMM_REL
    # the result is complex64-accurate, not exact: a plain copy would pass the
    # c128 bound only by accident
    assert np.abs(c64 - ref).max() > 0.0 or k == 1


def test_pt_tempo_c64_form():
    """PT-TEMPO + compute_dynamics in the precision form on the headline bath
    (alpha 0.5, w_c 5, T 0.2, dt 0.05, epsrel 1e-7) with dkmax 6, 24 steps."""
    _announce("pt_tempo c64 dkmax=6 N=24")
    system, bath, params = models.spin_boson(0.5, 5.0, 0.2, 0.05, 6, 1e-7)
    rho0 = op.spin_dm("z+")
    n = 24
    pt = oqupy.pt_tempo_compute(bath, 0.0, n * 0.05 + 1e-11, params,
                                backend_config={"precision": "c64"})
    dyn = oqupy.compute_dynamics(system=system, process_tensor=pt,
                                 initial_state=rho0)
    ref_states, ref_bonds, _ = models.oracle_pt_dynamics(system, bath, params,
                                                         rho0, n)
    sz = models.sigma_z_expectation(dyn.states)
    sz_ref = models.sigma_z_expectation(ref_states)
    dev = np.abs(sz - sz_ref).max()
    print(f"  c64 form: max |d<sz>| = {dev:.2e}; bonds "
          f"{list(pt.get_bond_dimensions()[1:-1])} vs {ref_bonds}")
    assert dev < C64_SIGMA_Z_TOL
    tr = np.real(np.trace(dyn.states, axis1=1, axis2=2))
    print(f"  c64 form: max |tr rho - 1| = {np.abs(tr - 1.0).max():.2e}")
    assert np.abs(tr - 1.0).max() < C64_TRACE_TOL


def test_tempo_c64_form():
    """TEMPO in the precision form on the rho_A golden case. TEMPO's
    truncation amplifies rounding-level differences (two backward-stable
    LAPACK drivers in the FP64 oracle already disagree on this case), so the
    bound is the form's TEMPO floor or 10x the drivers' spread of
    <sigma_z>, whichever is larger."""
    _announce("tempo c64 rho_A")
    system, bath, params, rho0, n, rho, dec = models.golden_model("rho_A")
    tempo = oqupy.Tempo(system, bath, params, rho0, start_time=0.0,
                        backend_config={"precision": "c64"})
    dyn = tempo.compute(end_time=n * params.dt + 1e-9 * params.dt)
    ref, _ = models.oracle_tempo(system, bath, params, rho0, n)
    import scipy.linalg as sl
    from oracle import chain
    orig = chain.truncated_svd

    def gesvd(mat, eps, spectra=None):
        u, s, vh = sl.svd(mat, full_matrices=False, lapack_driver="gesvd")
        k = chain.truncation_rank(s, eps)
        return u[:, :k], s[:k], vh[:k]
    chain.truncated_svd = gesvd
    try:
        alt, _ = models.oracle_tempo(system, bath, params, rho0, n)
    finally:
        chain.truncated_svd = orig
    sz_ref = models.sigma_z_expectation(ref)
    spread = np.abs(models.sigma_z_expectation(alt) - sz_ref).max()
    dev = np.abs(models.sigma_z_expectation(dyn.states) - sz_ref).max()
    print(f"  c64 TEMPO: max |d<sz>| = {dev:.2e}, LAPACK-driver spread of "
          f"<sz> {spread:.2e}")
    assert dev < max(C64_TEMPO_SIGMA_Z_TOL, 10 * spread)
    np.testing.assert_almost_equal(dyn.states[-1], rho, decimal=dec)


def test_unknown_precision_is_rejected():
    with pytest.raises(ValueError):
        engine.parse_backend_config({"precision": "bf16"})
