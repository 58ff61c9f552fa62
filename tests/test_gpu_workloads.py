"""Parity of the engine with the FP64 oracle on the benchmark workloads
(needs a GPU).

* TEMPO, spin-boson alpha 0.3, w_c 3, T 0, dkmax 40, epsrel 1e-6: the whole
  200-step run, states within the LAPACK-driver spread rule of the TEMPO
  golden tests (1e-10 until two LAPACK drivers disagree beyond 1e-12).
* PT-TEMPO, spin-boson alpha 0.5, w_c 5, T 0.2, dkmax 200, epsrel 1e-7
  (N = 1000): initialisation and the first 20 steps.
* PT-TEMPO, two qubits with a collective bath (Liouville dimension 16),
  dkmax 150, epsrel 1e-7 (N = 600): initialisation and the first 8 steps.
* PT-TEMPO, long memory alpha 0.6, dt 0.025, dkmax 400, epsrel 1e-8
  (N = 2000): initialisation and the first 12 steps.
* The batched form on the first 32 instances of the seeded sweep of baths
  (dkmax 50..300): 3 steps of every instance through pte_pt_batch_run.

For PT-TEMPO every truncated SVD is compared in order (the engine's SVD log
against the oracle's, the oracle continuing from the engine's MPS at every
step's start): the same matrix shape, the same kept rank (a different rank
only where the cut is rounding-ambiguous: reached by the spectrum moved by
1e-10 s_0, and then taken by both), and kept singular values within
1e-10 s_0; the bond dimensions afterwards are identical.
"""

import ctypes
import sys
import time

import numpy as np
import pytest

import oqupy_amd as oqupy
from oqupy_amd import operators as op
from oqupy_amd.backends import engine

from tests import models

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-10
SPECTRUM_TOL = 1e-10
LOG_VALUES = 50_000_000


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)
    sys.stdout.flush()


def _table(bath, params, count):
    infl = models.CachedInfluence(bath, params)
    return {dk: infl(dk) for dk in range(count)}


def _export(backend):
    """The engine's MPS as stored (for the oracle to continue from)."""
    n_sites = len(backend.bond_dimensions()) + 1
    return backend.step, [backend.get_site(i) for i in range(n_sites)]


def _pt_workload_parity(key, steps):
    """Engine and oracle side by side: both initialise from the same
    influence tables, then for every step the oracle continues from the
    engine's MPS at the step's start (so a step's SVDs are compared on the
    same input, not after the rounding of all earlier steps has grown
    through the truncations), and every SVD of the whole run is compared in
    order. Where the oracle's cut differs from the engine's only because it
    is rounding-ambiguous, the oracle keeps the engine's rank
    (``models.SteeredCut``)."""
    from oracle.pt_tempo import PtTempoOracle
    system, bath, params, spec = models.workload(key)
    dim, n, dkmax, epsrel = (spec["dim"], spec["nsteps"], spec["dkmax"],
                             spec["epsrel"])
    t0 = time.perf_counter()
    table = _table(bath, params, min(n, dkmax + 1))
    t1 = time.perf_counter()
    ptt = oqupy.PtTempo(bath, 0.0, n * params.dt + 1e-9 * params.dt,
                        params, backend_config={"svd_log": LOG_VALUES})
    assert ptt.num_steps == n
    backend = ptt.backend
    backend._influence = lambda dk: table.get(dk)
    ref = PtTempoOracle(dim, lambda dk: table.get(dk), np.ones(dim * dim),
                        n, dkmax, epsrel)
    ref.spectra = []
    steer = models.SteeredCut([], epsrel, SPECTRUM_TOL)
    t_engine = t_oracle = 0.0
    ta = time.perf_counter()
    backend.initialize()
    tb = time.perf_counter()
    steer.engine_log = backend.get_svd_log()
    with steer:
        ref.initialize()
    tc = time.perf_counter()
    t_engine += tb - ta
    t_oracle += tc - tb
    for _ in range(steps):
        state = _export(backend)
        ta = time.perf_counter()
        backend.compute_step()
        tb = time.perf_counter()
        steer.engine_log = backend.get_svd_log()
        ref.load_state(*state)
        with steer:
            ref.compute_step()
        tc = time.perf_counter()
        t_engine += tb - ta
        t_oracle += tc - tb
    log = backend.get_svd_log()
    worst = models.compare_svd_logs(log, ref.spectra, epsrel, SPECTRUM_TOL,
                                    steer.adopted)
    assert backend.bond_dimensions() == ref.bond_dimensions()
    print(f"  {key}: {len(log)} SVDs, max bond {max(ref.bond_dimensions())}, "
          f"max |ds|/s0 = {worst:.2e}, rounding-ambiguous cuts "
          f"{len(steer.adopted)}; influence {t1 - t0:.1f} s, engine "
          f"{t_engine:.1f} s, oracle {t_oracle:.1f} s", flush=True)
    backend.release()


def test_c2_headline_first_steps():
    _announce("workload c2: PT-TEMPO dkmax=200 epsrel=1e-7, init + 20 steps")
    _pt_workload_parity("c2", 20)


def test_c3_two_qubits_first_steps():
    _announce("workload c3: two qubits d^2=16 dkmax=150, init + 8 steps")
    _pt_workload_parity("c3", 8)


def test_c5_long_memory_first_steps():
    _announce("workload c5: dt=0.025 dkmax=400 epsrel=1e-8, init + 12 steps")
    _pt_workload_parity("c5", 12)


def test_c1_tempo_whole_run():
    """The whole 200-step TEMPO run of the short-memory workload."""
    _announce("workload c1: TEMPO dkmax=40 epsrel=1e-6, 200 steps")
    from tests.test_gpu_parity import _assert_within_spread, \
        _oracle_tempo_spread
    system, bath, params, spec = models.workload("c1")
    rho0 = op.spin_dm("z+")
    n = spec["nsteps"]
    tempo = oqupy.Tempo(system, bath, params, rho0, start_time=0.0)
    dyn = tempo.compute(end_time=n * params.dt + 1e-9 * params.dt)
    assert len(dyn.times) == n + 1
    ref, spread = _oracle_tempo_spread(system, bath, params, rho0, n)
    dsz = np.abs(models.sigma_z_expectation(dyn.states)
                 - models.sigma_z_expectation(ref)).max()
    print(f"  c1: max |d<sigma_z>| = {dsz:.2e}, LAPACK-driver spread "
          f"{spread.max():.2e}", flush=True)
    _assert_within_spread(np.asarray(dyn.states), ref, spread)


def test_c4_batched_instances():
    """32 instances of the seeded sweep (uneven memory cuts and bonds) stepped
    together by the batched form, each against its own oracle run."""
    _announce("workload c4: 32 instances x 3 steps, batched")
    from oracle.pt_tempo import PtTempoOracle
    from oqupy_amd.influence import device_influences
    count, steps, n = 32, 3, 500
    setups = [models.sweep_instance(i) for i in range(count)]
    baths = [b for b, _ in setups]
    params = [p for _, p in setups]
    dkmaxes = [p.dkmax for p in params]
    infls = device_influences(baths, params, dkmaxes)
    tables = [{dk: inf(dk) for dk in range(k + 1)}
              for inf, k in zip(infls, dkmaxes)]
    lib = engine.load_library()
    backends = []
    for i in range(count):
        ptt = oqupy.PtTempo(baths[i], 0.0, n * 0.05 + 1e-9, params[i],
                            backend_config={"svd_log": LOG_VALUES})
        tab = tables[i]
        ptt.backend._influence = lambda dk, tab=tab: tab.get(dk)
        ptt.backend.initialize()
        backends.append(ptt.backend)
    handles = (ctypes.c_void_p * count)(*[b.handle.value for b in backends])
    batch = ctypes.c_void_p()
    engine.check(lib.pte_pt_batch_create(handles, count, ctypes.byref(batch)))
    more = np.ones(count, dtype=np.int32)
    try:
        for _ in range(steps):
            engine.check(lib.pte_pt_batch_run(batch, 1, engine.iptr(more)))
    finally:
        lib.pte_pt_batch_destroy(batch)
    worst = 0.0
    for i, b in enumerate(backends):
        b.sync_step()
        assert b.step == 1 + steps
        ref = PtTempoOracle(2, lambda dk, t=tables[i]: t.get(dk), np.ones(4),
                            n, dkmaxes[i], 1e-7)
        ref.spectra = []
        ref.initialize()
        for _ in range(steps):
            ref.compute_step()
        worst = max(worst, models.compare_svd_logs(b.get_svd_log(),
                                                   ref.spectra, 1e-7,
                                                   SPECTRUM_TOL))
        assert b.bond_dimensions() == ref.bond_dimensions(), i
    print(f"  c4: {count} instances, max |ds|/s0 = {worst:.2e}", flush=True)
    for b in backends:
        b.release()
