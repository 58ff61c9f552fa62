"""Benchmark: PT-TEMPO time-steps/s of the spin-boson process tensor at
dkmax = 200 on one MI355X, and max |d<sigma_z>| against the FP64 oracle.

Workload (the headline configuration): H_S = (Omega/2) sigma_x with Omega = 1,
coupling sigma_z/2, Ohmic J(w) = 2 alpha w exp(-w/w_c) with alpha = 0.5,
w_c = 5, T = 0.2; dt = 0.05, dkmax = 200, epsrel = 1e-7, 1000 steps,
complex128. A "step" is one PT-TEMPO ``compute_step`` (zip-up of the
201-node influence MPO into the device-resident MPS plus the truncating
sweep). The influence tensors are computed on the host before the timed
region. The bond dimension of this workload grows by ~12 per step for the
first ~200 steps, so the run first goes untimed through steps 2..41 (max bond
~500 at step 40), then ``--warmup`` untimed steps, then ``--steps`` timed
steps between device synchronisations; the JSON line names the window.

The CPU baseline is the FP64 oracle (numpy/LAPACK) continuing from the
engine's MPS at the window start, with the BLAS threads set to the host's
cores and to one, for a bounded time; its rate is the share of the window's
first step (in algorithmic SVD flops, from the engine's own SVD log of that
step) it completes per second. The accuracy half of the metric comes from a
whole run the oracle finishes in seconds: the 200-step TEMPO workload
(dkmax 40, epsrel 1e-6), max |d<sigma_z>| over all 201 times.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6

# untimed steps that take each workload into its large-bond phase before the
# warm-up and the timed window (the headline run's bond dimension grows by
# ~12 per step and passes 500 at step 40)
PHASE_STEPS = {"c2": 40, "c3": 0, "c5": 0}

# HBM bytes per launch of the roofline kernel, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of the same command and window, scaled by
# what the counters read for a plain copy (keyed by config, precision, rank
# form, phase steps, warm-up, timed steps)
TRAFFIC = {
    # FETCH_SIZE 1366370.4 KiB x 2.000 + WRITE_SIZE 1104523.3 KiB x 1.000 per
    # Jacobi dispatch, mean over the window's 6580 dispatches
    ("c2", "c128", "exact", 40, 5, 20): 3929358419,
}
TRAFFIC_SOURCE = "profiles/hbm_traffic_svd_kernels_pt_tempo_dkmax200_steps47-66.md"

# PT-TEMPO workloads: headline spin-boson (c2), two qubits with a collective
# bath (c3, Liouville dimension 16), and the long-memory, tight-tolerance run
# of the learned form (c5)
PT_CONFIGS = {
    "c2": dict(name="spin-boson PT-TEMPO", dim=2, alpha=0.5, cutoff=5.0,
               temp=0.2, dt=0.05, dkmax=200, epsrel=1e-7, nsteps=1000),
    "c3": dict(name="two-qubit collective-bath PT-TEMPO", dim=4, alpha=0.2,
               cutoff=4.0, temp=0.5, dt=0.05, dkmax=150, epsrel=1e-7,
               nsteps=600),
    "c5": dict(name="spin-boson PT-TEMPO (long memory)", dim=2, alpha=0.6,
               cutoff=5.0, temp=0.1, dt=0.025, dkmax=400, epsrel=1e-8,
               nsteps=2000),
}
CFG = dict(PT_CONFIGS["c2"])
ALPHA, CUTOFF, TEMP, DT, DKMAX, EPSREL, NSTEPS = 0.5, 5.0, 0.2, 0.05, 200, \
    1e-7, 1000


def select_config(key):
    """Make `key` the PT-TEMPO workload of this run."""
    global ALPHA, CUTOFF, TEMP, DT, DKMAX, EPSREL, NSTEPS
    CFG.clear()
    CFG.update(PT_CONFIGS[key])
    ALPHA, CUTOFF, TEMP, DT = CFG["alpha"], CFG["cutoff"], CFG["temp"], CFG["dt"]
    DKMAX, EPSREL, NSTEPS = CFG["dkmax"], CFG["epsrel"], CFG["nsteps"]


def coupling():
    """S = sigma_z / 2 (one spin) or (sigma_z x 1 + 1 x sigma_z) / 2."""
    import numpy as np
    from oqupy_amd import operators as op
    if CFG["dim"] == 2:
        return 0.5 * op.sigma("z")
    eye = np.eye(2)
    return 0.5 * (np.kron(op.sigma("z"), eye) + np.kron(eye, op.sigma("z")))


def build_model():
    import numpy as np
    from oqupy_amd.bath import Bath
    from oqupy_amd.bath_correlations import PowerLawSD
    from oqupy_amd.tempo import TempoParameters
    corr = PowerLawSD(alpha=ALPHA, zeta=1.0, cutoff=CUTOFF,
                      cutoff_type="exponential", temperature=TEMP)
    bath = Bath(coupling(), corr)
    params = TempoParameters(dt=DT, dkmax=DKMAX, epsrel=EPSREL)
    return bath, params, np


def influence_table(bath, params, count=None):
    from oqupy_amd.tempo import influence_matrix
    corr = bath.correlations
    count = DKMAX + 1 if count is None else count
    return {dk: influence_matrix(dk, params, corr, bath.coupling_acomm,
                                 bath.coupling_comm)
            for dk in range(count)}


def train_rank_predictor():
    """The rank predictor trained on the engine's own traces of PT-TEMPO runs
    of two other baths (alpha 0.3, w_c 3, T 0 and alpha 0.2, w_c 4, T 0.5;
    dkmax 20, epsrel 1e-7, 40 steps); returns (predictor, summary)."""
    import numpy as np
    import oqupy_amd as oqupy
    from oqupy_amd import operators as op
    from oqupy_amd.rank_predictor import RankPredictor
    t0 = time.perf_counter()
    records = []
    for alpha, wc, temp in ((0.3, 3.0, 0.0), (0.2, 4.0, 0.5)):
        corr = oqupy.PowerLawSD(alpha=alpha, zeta=1.0, cutoff=wc,
                                cutoff_type="exponential", temperature=temp)
        bath = oqupy.Bath(0.5 * op.sigma("z"), corr)
        params = oqupy.TempoParameters(dt=DT, dkmax=20, epsrel=EPSREL)
        ptt = oqupy.PtTempo(bath, 0.0, 40 * DT + 1e-9, params,
                            backend_config={"rank_trace": 100000})
        ptt.compute()
        records.append(ptt.backend.rank_trace())
    records = np.concatenate(records)
    pred = RankPredictor(seed=0)
    losses = pred.fit(records, epochs=60, batch=256, lr=1e-3, seed=1)
    return pred, {"records": int(len(records)),
                  "loss_first_last": [float(losses[0]), float(losses[-1])],
                  "seconds": round(time.perf_counter() - t0, 2)}


def svd_flops(m, n):
    """Algorithmic flops of a thin complex SVD of an m x n matrix: the
    R-SVD count of Golub & Van Loan, 4 (6 M N^2 + 20 N^3) with M = max and
    N = min (complex arithmetic = 4 real)."""
    big, small = max(m, n), min(m, n)
    return 4.0 * (6.0 * big * small * small + 20.0 * small ** 3)


def gpu_run(warmup, steps, table, precision="c128", rank_predictor=None,
            export_state=False, rank_theta=None, progress=False):
    """The engine on the workload: initialise, `warmup` untimed steps, then
    `steps` timed steps. Returns a dict: elapsed seconds, Jacobi kernel
    statistics of the window, per-step statistics, the SVD log of the window
    (m, n, kept spectrum per SVD) and, with export_state, the MPS sites at
    the window start (the CPU leg continues from them)."""
    import torch
    import oqupy_amd as oqupy
    bath, params, np = build_model()
    cfg = {"backend": "mi355x", "precision": precision}
    if rank_predictor is not None:
        cfg["rank_predictor"] = rank_predictor
        if rank_theta is not None:
            cfg["rank_theta"] = rank_theta
    ptt = oqupy.PtTempo(bath, 0.0, NSTEPS * DT + 1e-9, params,
                        backend_config=cfg)
    assert ptt.num_steps == NSTEPS
    ptt._influence = lambda dk: table.get(dk)
    backend = ptt.backend
    backend._influence = ptt._influence
    backend.initialize()
    for i in range(warmup):
        backend.compute_step()
        if (i + 1) % 10 == 0:
            print(f"[bench] step {backend.step} max bond "
                  f"{backend.step_stats[-1][2]}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    state = None
    if export_state:
        n_sites = len(backend.bond_dimensions()) + 1
        state = (backend.step, [backend.get_site(i) for i in range(n_sites)])
    # the SVD log (shape and kept spectrum of every truncated SVD) gives the
    # algorithmic work of the window; it costs one small device-to-host copy
    # per SVD
    backend.svd_log(200_000_000)
    before = backend.svd_kernel_stats()
    t0 = time.perf_counter()
    for i in range(steps):
        backend.compute_step()
        if progress:
            # a line per timed step for long profiled runs (a few microseconds
            # against a step of about two seconds)
            print(f"[bench] timed step {i + 1}/{steps}", file=sys.stderr,
                  flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _progress(f"timed window: {steps} steps in {elapsed:.1f} s")
    after = backend.svd_kernel_stats()
    window = backend.step_stats[-steps:]
    svd = {k: after[k] - before[k] for k in after}
    rank = backend.rank_stats() if rank_predictor is not None else None
    log = backend.get_svd_log()
    backend.release()
    return {"elapsed": elapsed, "svd": svd, "window": window,
            "last_step": backend.step, "rank": rank, "log": log,
            "state": state}


def blas_threads():
    """Threads of the BLAS/LAPACK library numpy uses (the oracle's cores)."""
    try:
        from threadpoolctl import threadpool_info
        n = [d.get("num_threads", 0) for d in threadpool_info()
             if d.get("user_api") == "blas"]
        if n:
            return int(max(n))
    except Exception:
        pass
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))


class _BudgetSpent(Exception):
    pass


class _Deadline(list):
    """The oracle's SVD log that ends its run once the time budget is spent
    (checked after every completed SVD)."""

    def __init__(self, seconds):
        super().__init__()
        self.t_end = time.perf_counter() + seconds

    def append(self, item):
        super().append(item)
        if time.perf_counter() > self.t_end:
            raise _BudgetSpent()


def cpu_baseline(state, table, step_flops, budget_s, threads):
    """The FP64 oracle continuing from the engine's state at the window
    start (the same MPS sites), with the BLAS/LAPACK library limited to
    `threads` threads, for about `budget_s` seconds of the window's first
    step. Its rate in steps/s is the share of that step's algorithmic SVD
    flops completed (per SVD, from the engine's log of the same step) per
    second. Returns (steps/s, SVDs completed, flop share completed)."""
    from threadpoolctl import threadpool_limits
    from oracle.pt_tempo import PtTempoOracle
    import numpy as np
    d = CFG["dim"]
    step, sites = state
    pt = PtTempoOracle(d, lambda dk: table.get(dk), np.ones(d * d), NSTEPS,
                       DKMAX, EPSREL)
    pt.load_state(step, sites)
    with threadpool_limits(limits=threads, user_api="blas"):
        pt.spectra = _Deadline(budget_s)
        t0 = time.perf_counter()
        try:
            pt.compute_step()
        except _BudgetSpent:
            pass
        elapsed = time.perf_counter() - t0
    done = sum(svd_flops(m, n) for m, n, _ in pt.spectra)
    share = done / step_flops
    return share / elapsed, len(pt.spectra), share


C4_SEED, C4_N, C4_DT, C4_EPSREL = 0, 500, 0.05, 1e-7


def c4_instance(i):
    """Instance i of the seeded sweep of spin-boson baths: alpha ~ U[0.05, 0.8],
    w_c ~ U[1, 10], T ~ U[0, 1], dkmax ~ U{50..300}, dt 0.05, epsrel 1e-7."""
    import numpy as np
    rng = np.random.default_rng(C4_SEED)
    alphas = rng.uniform(0.05, 0.8, 256)
    wcs = rng.uniform(1.0, 10.0, 256)
    temps = rng.uniform(0.0, 1.0, 256)
    dkmaxs = rng.integers(50, 301, 256)
    return float(alphas[i]), float(wcs[i]), float(temps[i]), int(dkmaxs[i])


def c4_table(i):
    from oqupy_amd import operators as op
    from oqupy_amd.bath import Bath
    from oqupy_amd.bath_correlations import PowerLawSD
    from oqupy_amd.tempo import TempoParameters, influence_matrix
    a, wc, t, k = c4_instance(i)
    corr = PowerLawSD(alpha=a, zeta=1.0, cutoff=wc, cutoff_type="exponential",
                      temperature=t)
    bath = Bath(0.5 * op.sigma("z"), corr)
    params = TempoParameters(dt=C4_DT, dkmax=k, epsrel=C4_EPSREL)
    return {dk: influence_matrix(dk, params, bath.correlations,
                                 bath.coupling_acomm, bath.coupling_comm)
            for dk in range(k + 1)}


def c4_tables_device(count):
    """Influence tables of the first `count` instances from eta evaluated on
    the GPU, all baths in one launch (the host QUADPACK path takes 10-30 s
    per instance)."""
    from oqupy_amd import operators as op
    from oqupy_amd.bath import Bath
    from oqupy_amd.bath_correlations import PowerLawSD
    from oqupy_amd.influence import device_influences
    from oqupy_amd.tempo import TempoParameters
    baths, params, dkmaxes = [], [], []
    for i in range(count):
        a, wc, t, k = c4_instance(i)
        corr = PowerLawSD(alpha=a, zeta=1.0, cutoff=wc,
                          cutoff_type="exponential", temperature=t)
        baths.append(Bath(0.5 * op.sigma("z"), corr))
        params.append(TempoParameters(dt=C4_DT, dkmax=k, epsrel=C4_EPSREL))
        dkmaxes.append(k)
    infls = device_influences(baths, params, dkmaxes)
    return [{dk: inf(dk) for dk in range(k + 1)}
            for inf, k in zip(infls, dkmaxes)]


def c4_engines(count, tables):
    import oqupy_amd as oqupy
    from oqupy_amd import operators as op
    out = []
    for i in range(count):
        a, wc, t, k = c4_instance(i)
        corr = oqupy.PowerLawSD(alpha=a, zeta=1.0, cutoff=wc,
                                cutoff_type="exponential", temperature=t)
        bath = oqupy.Bath(0.5 * op.sigma("z"), corr)
        params = oqupy.TempoParameters(dt=C4_DT, dkmax=k, epsrel=C4_EPSREL)
        ptt = oqupy.PtTempo(bath, 0.0, C4_N * C4_DT + 1e-9, params,
                            backend_config={})
        tab = tables[i]
        ptt.backend._influence = lambda dk, tab=tab: tab.get(dk)
        ptt.backend.initialize()
        out.append(ptt.backend)
    return out


def batch_main(args):
    """Batched form: the first --batch instances of the seeded sweep, one
    warm-up step each, then --steps steps each, all engines stepped together
    through batched SVD launches; the same instances stepped one by one on
    the GPU are timed next to it."""
    import ctypes
    from concurrent.futures import ProcessPoolExecutor
    import numpy as np
    import torch
    from oqupy_amd.backends import engine as eng
    count = args.batch
    t0 = time.perf_counter()
    if args.host_influence:
        with ProcessPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
            tables = list(ex.map(c4_table, range(count)))
    else:
        tables = c4_tables_device(count)
    precompute_s = time.perf_counter() - t0
    print(f"influence tables of {count} instances: {precompute_s:.2f} s",
          file=sys.stderr, flush=True)
    lib = eng.load_library()

    def note(msg):
        print(f"[batch {count}] {msg} ({time.perf_counter() - t0:.1f} s)",
              file=sys.stderr, flush=True)
    backends = c4_engines(count, tables)
    note("engines initialised")
    handles = (ctypes.c_void_p * count)(*[b.handle.value for b in backends])
    batch = ctypes.c_void_p()
    eng.check(lib.pte_pt_batch_create(handles, count, ctypes.byref(batch)))
    more = np.ones(count, dtype=np.int32)
    eng.check(lib.pte_pt_batch_run(batch, 1, eng.iptr(more)))
    torch.cuda.synchronize()
    note("batched warm-up step done")
    t1 = time.perf_counter()
    for _ in range(args.steps):
        eng.check(lib.pte_pt_batch_run(batch, 1, eng.iptr(more)))
        note("batched step done")
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t1
    counts = [ctypes.c_long() for _ in range(3)]
    times = [ctypes.c_double() for _ in range(3)]
    eng.check(lib.pte_pt_batch_stats(batch, *[ctypes.byref(c) for c in counts],
                                     *[ctypes.byref(t) for t in times]))
    lib.pte_pt_batch_destroy(batch)
    for b in backends:
        b.sync_step()
    bonds = [max(b.bond_dimensions()) for b in backends]
    del backends
    single = c4_engines(count, tables)
    for b in single:
        b.compute_step()
    torch.cuda.synchronize()
    note("one-by-one engines ready")
    t1 = time.perf_counter()
    for _ in range(args.steps):
        for b in single:
            b.compute_step()
        note("one-by-one step done")
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t1
    value = count * args.steps / t_batch
    print(json.dumps({
        "metric": "batched PT-TEMPO instance-steps/s (sum over instances)",
        "value": round(value, 4), "unit": "instance-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1,
        "ms_per_step": round(1e3 * t_batch / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"first {count} instances of the seeded sweep "
                               f"(alpha~U[0.05,0.8], wc~U[1,10], T~U[0,1], "
                               f"dkmax~U{{50..300}}, dt=0.05, epsrel=1e-7, "
                               f"N=500, seed 0), steps 2..{1 + args.steps}",
                   "max_bond_in_window": int(max(bonds)),
                   "one_by_one_instance_steps_per_s": round(
                       count * args.steps / t_single, 4),
                   "batch_speedup": round(t_single / t_batch, 3),
                   "batches": counts[0].value,
                   "svds_per_batch": round(counts[1].value
                                           / max(counts[0].value, 1), 2),
                   "cooperative_launches": counts[2].value,
                   "serve_s": round(times[0].value / 1e3, 3),
                   "batched_qr_s": round(times[1].value / 1e3, 3),
                   "batched_jacobi_s": round(times[2].value / 1e3, 3),
                   "influence_precompute_s": round(precompute_s, 2),
                   "influence": "host QUADPACK" if args.host_influence
                                else "eta on the GPU (pte_eta_batch)"},
        "roofline": None, "cpu_baseline": None}))


def _progress(msg):
    """A progress line on stderr between the benchmark's phases (never
    inside the timed window)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class _OtherDriver:
    """While active, the oracle's truncated SVD uses LAPACK's gesvd instead
    of gesdd: the spread between two backward-stable drivers is how far the
    FP64 reference itself defines a result whose truncations sit at
    near-ties (loose epsrel)."""

    def __enter__(self):
        import scipy.linalg as sl
        from oracle import chain
        self._chain, self._orig = chain, chain.truncated_svd

        def gesvd(mat, eps, spectra=None):
            u, s, vh = sl.svd(mat, full_matrices=False, lapack_driver="gesvd")
            if spectra is not None:
                spectra.append((mat.shape[0], mat.shape[1], s.copy()))
            k = chain.truncation_rank(s, eps)
            return u[:, :k], s[:k], vh[:k]
        chain.truncated_svd = gesvd
        return self

    def __exit__(self, *exc):
        self._chain.truncated_svd = self._orig
        return False


def tempo_c1(precision="c128", cpu=True, spread=False):
    """TEMPO on the short-memory spin-boson workload (H_S = sigma_x / 2,
    coupling sigma_z / 2, Ohmic alpha 0.3, w_c 3, T 0, rho_0 = |up><up|,
    dt 0.05, dkmax 40, epsrel 1e-6, 200 steps) on the engine and, with
    `cpu`, on the FP64 oracle with the same influences. Returns (GPU
    seconds for steps 2..200, oracle seconds for the same steps or None,
    max |d<sigma_z>| over all 201 times or None) and, with `spread`, the
    oracle's own spread between two LAPACK drivers."""
    import numpy as np
    import torch
    import oqupy_amd as oqupy
    from oqupy_amd import operators as op
    from oqupy_amd.tempo import influence_matrix
    from oracle.tempo import TempoOracle
    n, dkmax = 200, 40
    corr = oqupy.PowerLawSD(alpha=0.3, zeta=1.0, cutoff=3.0,
                            cutoff_type="exponential", temperature=0.0)
    bath = oqupy.Bath(0.5 * op.sigma("z"), corr)
    system = oqupy.System(0.5 * op.sigma("x"))
    params = oqupy.TempoParameters(dt=0.05, dkmax=dkmax, epsrel=1e-6)
    rho0 = op.spin_dm("up")
    table = {dk: influence_matrix(dk, params, bath.correlations,
                                  bath.coupling_acomm, bath.coupling_comm)
             for dk in range(dkmax + 1)}
    tempo = oqupy.Tempo(system, bath, params, rho0, 0.0,
                        backend_config={"precision": precision})
    tempo._backend_instance._influence = lambda dk: table.get(dk)
    tempo.compute(end_time=0.05 * 1 + 1e-9)          # initialise + step 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dyn = tempo.compute(end_time=0.05 * n + 1e-9)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tempo._backend_instance.release()
    if not cpu:
        return elapsed, None, None
    props = system.get_propagators(0.05, 0.0, params.subdiv_limit,
                                   params.liouvillian_epsrel)
    sz = op.sigma("z")

    def oracle_run():
        ref = TempoOracle(rho0, lambda dk: table.get(dk), bath.unitary_transform,
                          props, np.ones(4), np.ones(4), dkmax, 1e-6)
        _, st = ref.initialize()
        states = [st.reshape(2, 2)]
        _, st = ref.compute_step()
        states.append(st.reshape(2, 2))
        t0 = time.perf_counter()
        for _ in range(n - 1):
            _, st = ref.compute_step()
            states.append(st.reshape(2, 2))
        return (time.perf_counter() - t0,
                np.real(np.einsum("ij,tji->t", sz, np.asarray(states))))
    _progress("TEMPO oracle")
    cpu_elapsed, sz_ref = oracle_run()
    dsz = np.abs(np.real(np.einsum("ij,tji->t", sz, np.asarray(dyn.states)))
                 - sz_ref)
    if not spread:
        return elapsed, cpu_elapsed, float(dsz.max())
    _progress("TEMPO oracle, second LAPACK driver")
    with _OtherDriver():
        _, sz_alt = oracle_run()
    return elapsed, cpu_elapsed, float(dsz.max()), float(np.abs(sz_alt - sz_ref).max())


def pt_accuracy(precision="c128", rank_predictor=None, rank_theta=None):
    """max |d<sigma_z>| of a whole PT-TEMPO + compute_dynamics run in the
    given form against the FP64 oracle fed the same influences: the
    short-memory spin-boson bath (alpha 0.3, w_c 3, T 0), dt 0.05, dkmax 20,
    epsrel 1e-6, 80 steps (the oracle finishes it in seconds). Returns the
    deviation, the oracle's largest bond and the oracle's spread between
    two LAPACK drivers."""
    import numpy as np
    import oqupy_amd as oqupy
    from oqupy_amd import operators as op
    from oqupy_amd.tempo import influence_matrix
    from oracle.dynamics import compute_dynamics
    from oracle.pt_tempo import PtTempoOracle, compute_caps
    n, dkmax, eps = 80, 20, 1e-6
    corr = oqupy.PowerLawSD(alpha=0.3, zeta=1.0, cutoff=3.0,
                            cutoff_type="exponential", temperature=0.0)
    bath = oqupy.Bath(0.5 * op.sigma("z"), corr)
    system = oqupy.System(0.5 * op.sigma("x"))
    params = oqupy.TempoParameters(dt=0.05, dkmax=dkmax, epsrel=eps)
    rho0 = op.spin_dm("up")
    table = {dk: influence_matrix(dk, params, bath.correlations,
                                  bath.coupling_acomm, bath.coupling_comm)
             for dk in range(dkmax + 1)}
    cfg = {"precision": precision}
    if rank_predictor is not None:
        cfg["rank_predictor"] = rank_predictor
        if rank_theta is not None:
            cfg["rank_theta"] = rank_theta
    ptt = oqupy.PtTempo(bath, 0.0, n * 0.05 + 1e-9, params, backend_config=cfg)
    ptt._influence = lambda dk: table.get(dk)
    ptt.backend._influence = ptt._influence
    ptt.compute()
    pt = ptt.get_process_tensor()
    dyn = oqupy.compute_dynamics(system=system, process_tensor=pt,
                                 initial_state=rho0)
    sz = op.sigma("z")
    props = system.get_propagators(0.05, 0.0, None, None)

    def oracle_run():
        ref = PtTempoOracle(2, lambda dk: table.get(dk), np.ones(4), n, dkmax, eps)
        ref.compute()
        tens = ref.mpo_tensors()
        states = compute_dynamics(props, rho0, n, [tens], [compute_caps(tens, 2)])
        return (np.real(np.einsum("ij,tji->t", sz, np.asarray(states))),
                max(ref.bond_dimensions()))
    _progress("PT-TEMPO oracle")
    sz_ref, bond = oracle_run()
    dsz = np.abs(np.real(np.einsum("ij,tji->t", sz, np.asarray(dyn.states)))
                 - sz_ref)
    _progress("PT-TEMPO oracle, second LAPACK driver")
    with _OtherDriver():
        sz_alt, _ = oracle_run()
    return float(dsz.max()), bond, float(np.abs(sz_alt - sz_ref).max())


def accuracy(precision="c128", rank_predictor=None, rank_theta=None):
    """The accuracy half of the metric on whole runs that the oracle
    finishes in seconds: max |d<sigma_z>| against the FP64 oracle of the
    200-step TEMPO workload (plain form) and of a 80-step PT-TEMPO +
    compute_dynamics run in the benchmarked form, each beside the oracle's
    own spread between two LAPACK drivers (at epsrel 1e-6 truncations at
    near-ties make both runs differ from the oracle by more than rounding:
    the drivers' spread is the scale against which to read the deviation)."""
    elapsed, cpu_elapsed, dsz, spread = tempo_c1(spread=True)
    pt_dsz, pt_bond, pt_spread = pt_accuracy(precision, rank_predictor, rank_theta)
    return {"tempo": {"workload": "spin-boson TEMPO alpha=0.3 wc=3.0 T=0 "
                                  "dt=0.05 dkmax=40 epsrel=1e-6 N=200 (whole "
                                  "run)",
                      "max_abs_dsigma_z": dsz,
                      "oracle_lapack_driver_spread": spread,
                      "gpu_steps_per_s": round(199 / elapsed, 3),
                      "oracle_steps_per_s": round(199 / cpu_elapsed, 3)},
            "pt_tempo": {"workload": "spin-boson PT-TEMPO alpha=0.3 wc=3.0 "
                                     "T=0 dt=0.05 dkmax=20 epsrel=1e-6 N=80 "
                                     "+ compute_dynamics (whole run), "
                                     "benchmarked form",
                         "max_abs_dsigma_z": pt_dsz,
                         "oracle_lapack_driver_spread": pt_spread,
                         "max_bond": pt_bond}}


def tempo_main(args):
    """The TEMPO workload on its own line: the whole step loop timed
    (influences precomputed), the FP64 oracle on the same steps on the host
    cores for the baseline and the accuracy."""
    steps = 199
    elapsed, cpu_elapsed, dsz = tempo_c1(args.precision)
    print(json.dumps({
        "metric": "TEMPO time-steps/s (dkmax=40); max |d<sigma_z>| vs FP64 oracle",
        "value": round(steps / elapsed, 4), "unit": "time-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": round(1e3 * elapsed / steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {"workload": "spin-boson TEMPO alpha=0.3 wc=3.0 T=0 dt=0.05 "
                               "dkmax=40 epsrel=1e-6 N=200, steps 2..200",
                   "max_abs_dsigma_z": dsz},
        "roofline": None,
        "cpu_baseline": {"value": round(steps / cpu_elapsed, 4),
                         "unit": "time-steps/s", "cores": blas_threads(),
                         "kind": "port",
                         "sample": "FP64 oracle on steps 2..200 of the same run"}}))


def trace_summary(directory, kernel, last):
    """Summary of a rocprofv3 output directory for one kernel: its dispatches
    in the kernel trace (all, and the last `last`, the benchmark's timed
    window) and, from a counter collection, the mean of every counter per
    dispatch over the same last dispatches. Prints one JSON object."""
    import csv
    import glob
    out = {"kernel": kernel, "directory": directory}

    def rows(pattern):
        found = []
        for path in sorted(glob.glob(os.path.join(directory, "**", pattern),
                                     recursive=True)):
            with open(path, newline="") as fh:
                found += list(csv.DictReader(fh))
        return found
    trace = [r for r in rows("*kernel_trace.csv")
             if kernel in r.get("Kernel_Name", "")]
    if trace:
        trace.sort(key=lambda r: int(r["Start_Timestamp"]))
        dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
               for r in trace]
        win = dur[-last:] if last > 0 else dur
        out["dispatches"] = len(dur)
        out["avg_ms_all"] = sum(dur) / len(dur)
        out["window_dispatches"] = len(win)
        out["avg_ms_window"] = sum(win) / len(win)
        # GPU busy time from the window's first dispatch of this kernel to
        # the last kernel of the run (union of all kernels' intervals)
        t_start = int(trace[-len(win)]["Start_Timestamp"])
        spans = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
                       for r in rows("*kernel_trace.csv")
                       if int(r["Start_Timestamp"]) >= t_start)
        busy, last_end = 0, t_start
        for s0, e0 in spans:
            busy += max(0, e0 - max(s0, last_end))
            last_end = max(last_end, e0)
        out["window_wall_ms"] = (last_end - t_start) / 1e6
        out["window_busy_frac"] = busy / max(last_end - t_start, 1)
    counters = [r for r in rows("*counter_collection.csv")
                if kernel in r.get("Kernel_Name", "")]
    if counters:
        per = {}
        for r in counters:
            per.setdefault(r["Counter_Name"], []).append(
                (int(r.get("Dispatch_Id", 0) or 0), float(r["Counter_Value"])))
        for name, vals in per.items():
            vals.sort()
            sel = [v for _, v in vals][-last:] if last > 0 else \
                [v for _, v in vals]
            out[f"{name}_mean_window"] = sum(sel) / len(sel)
            out[f"{name}_dispatches"] = len(vals)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--phase-steps", type=int, default=-1,
                    help="untimed steps before the warm-up that take the run "
                         "into its large-bond phase (default: 40 for c2, 0 "
                         "otherwise)")
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of oracle work for the CPU baseline (half "
                         "of it again for the one-thread leg)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="BLAS threads of the CPU baseline (default: the "
                         "host's cores, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--progress", action="store_true",
                    help="print a line after every timed step (profiled runs)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the whole-run TEMPO accuracy check")
    ap.add_argument("--config", choices=("c2", "c1", "c3", "c5"), default="c2",
                    help="c1: the TEMPO configuration; c3: two qubits with a "
                         "collective bath (Liouville dimension 16); c5: the "
                         "long-memory run (dt 0.025, dkmax 400, epsrel 1e-8); "
                         "default c2, the headline PT-TEMPO run")
    ap.add_argument("--batch", type=int, default=0,
                    help="time the batched form on this many instances of the "
                         "seeded sweep of baths instead of the headline run")
    ap.add_argument("--host-influence", action="store_true",
                    help="--batch: influence tables from host QUADPACK eta "
                         "instead of the GPU eta kernel")
    ap.add_argument("--precision", choices=("c128", "c64"), default="c128",
                    help="c64: contractions in complex64 on the matrix cores "
                         "(FP32 accumulation), SVD and cut in FP64")
    ap.add_argument("--rank", choices=("exact", "mlp", "previous", "adaptive"),
                    default="exact",
                    help="learned truncation-rank form: cap every pivoted QR "
                         "at the MLP's prediction (trained first on traces of "
                         "two other baths), at the bond's previous step count, "
                         "or stop it adaptively")
    ap.add_argument("--rank-theta", type=float, default=None,
                    help="acceptance threshold of a capped pivoted QR, as a "
                         "fraction of epsrel |beta_0| (learned forms only; "
                         "default 0.02)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 4 times (the "
                         "plain copy kernel that calibrates the HBM byte "
                         "counters of a profile) and exit")
    ap.add_argument("--trace-summary", metavar="DIR", default=None,
                    help="summarise a rocprofv3 output directory for --kernel "
                         "over its last --last dispatches and exit")
    ap.add_argument("--kernel", default="pte::jacobi")
    ap.add_argument("--last", type=int, default=0)
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary, args.kernel, args.last)
        return
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n, "launches": 4}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; a "
                          "single PT-TEMPO instance is a sequential chain of "
                          "SVDs and is not sharded across GPUs"}))
        sys.exit(2)
    if args.config == "c1":
        tempo_main(args)
        return
    if args.batch > 0:
        if args.batch > 256:
            raise SystemExit("--batch takes at most 256 instances")
        batch_main(args)
        return

    if args.config in PT_CONFIGS:
        select_config(args.config)
    bath, params, np = build_model()
    t0 = time.perf_counter()
    # influence tensors dk = 0 .. dkmax (the MPO), computed before the timed
    # region like the reference's own timing (its progress timer starts after
    # initialize)
    table = influence_table(bath, params)
    precompute_s = time.perf_counter() - t0
    phase = args.phase_steps if args.phase_steps >= 0 else PHASE_STEPS[
        args.config]
    warm = phase + args.warmup
    first = warm + 2

    predictor, training = None, None
    if args.rank == "mlp":
        predictor, training = train_rank_predictor()
    elif args.rank != "exact":
        predictor = args.rank
    run = gpu_run(warm, args.steps, table, args.precision, predictor,
                  export_state=not args.no_cpu, rank_theta=args.rank_theta,
                  progress=args.progress)
    elapsed, svd, window, log = run["elapsed"], run["svd"], run["window"], \
        run["log"]
    value = args.steps / elapsed
    per_step = [w[0] for w in window]
    first_step_log = log[:per_step[0]]
    step_flops = [sum(svd_flops(m, n) for m, n, _ in
                      log[sum(per_step[:i]):sum(per_step[:i + 1])])
                  for i in range(len(per_step))]
    svd_avg_ms = svd["ms"] / max(svd["count"], 1)
    achieved = svd["flops"] / max(svd["ms"], 1e-12) / 1e9   # TFLOP/s
    roofline = {
        "kernel": "pte::jacobi (jacobi_tile_kernel and jacobi_kernel: blocked "
                  "one-sided Jacobi SVD, FP64)",
        "bound": "mfma", "achieved": round(achieved, 4),
        "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 6),
        "traffic": TRAFFIC.get((args.config, args.precision, args.rank,
                                phase, args.warmup, args.steps)),
        "launches": svd["count"], "avg_launch_ms": round(svd_avg_ms, 4),
        "algorithmic_flops_per_launch": svd["flops"] / max(svd["count"], 1),
        "timing": "HIP events around every launch on the engine's stream",
    }
    if roofline["traffic"] is not None:
        roofline["traffic_source"] = TRAFFIC_SOURCE
    cpu = None
    if not args.no_cpu:
        cores = os.cpu_count() if args.cpu_threads <= 0 else args.cpu_threads
        cores = min(cores, 16)
        legs = []
        for threads, budget in ((cores, args.cpu_budget),
                                (1, args.cpu_budget / 2)):
            _progress(f"cpu baseline, {threads} BLAS threads")
            rate, n_done, share = cpu_baseline(run["state"], table,
                                               step_flops[0], budget, threads)
            legs.append({"value": round(rate, 6), "cores": threads,
                         "svds": n_done, "flop_share": round(share, 4),
                         "budget_s": budget})
        # the faster of the two thread counts is the baseline
        best = max(legs, key=lambda leg: leg["value"])
        cpu = {"value": best["value"], "unit": "time-steps/s",
               "cores": best["cores"], "kind": "port",
               "sample": (f"FP64 oracle (numpy/LAPACK zgesdd) continuing "
                          f"from the engine's MPS at the window start, step "
                          f"{first}: its first {best['svds']} truncated SVDs "
                          f"({100 * best['flop_share']:.1f} % of that step's "
                          f"algorithmic SVD flops, from the engine's log of "
                          f"the same step) in ~{best['budget_s']:.0f} s; "
                          f"steps/s = flop share / seconds; the faster of "
                          f"{cores} BLAS threads and 1"),
               "legs": legs}
    out = {
        "metric": f"PT-TEMPO time-steps/s at dkmax={DKMAX}",
        "value": round(value, 6), "unit": "time-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {
            "workload": (f"{CFG['name']} alpha={ALPHA} wc={CUTOFF} "
                         f"T={TEMP} dt={DT} dkmax={DKMAX} epsrel={EPSREL} "
                         f"N={NSTEPS}; steps 2..{phase + 1} untimed to reach "
                         f"the large-bond phase, {args.warmup} untimed "
                         f"warm-up steps, timed steps {first}.."
                         f"{first + args.steps - 1}"),
            "precision": ("plain complex128" if args.precision == "c128" else
                          "complex64 contractions (FP32 accumulation), "
                          "FP64 SVD and cut"),
            "max_bond_at_window_start": int(max(
                max(1, len(sp)) for _, _, sp in first_step_log)),
            "max_bond_in_window": max(w[2] for w in window),
            "svds_per_step": round(sum(per_step) / len(per_step), 1),
            "largest_svd_in_window": max(((m, n) for m, n, _ in log),
                                         key=lambda x: x[0] * x[1]),
            "algorithmic_svd_tflop_per_step": round(
                sum(step_flops) / len(step_flops) / 1e12, 3),
            "jacobi_sweeps_per_svd": round(
                sum(w[1] for w in window) / max(1, sum(per_step)), 2),
            "influence_precompute_s": round(precompute_s, 2),
            "rank_form": args.rank,
            "rank_theta": args.rank_theta,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if run["rank"] is not None:
        out["config"]["rank_stats_whole_run"] = run["rank"]
        out["config"]["rank_predictor_training"] = training
    if args.config == "c2" and not args.no_accuracy:
        _progress("accuracy runs")
        out["accuracy"] = accuracy(args.precision, predictor, args.rank_theta)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
