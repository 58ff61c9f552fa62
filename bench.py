"""Benchmark: PT-TEMPO time-steps/s of the spin-boson process tensor at
dkmax = 200 on one MI355X.

Workload (the headline configuration): H_S = (Omega/2) sigma_x with Omega = 1,
coupling sigma_z/2, Ohmic J(w) = 2 alpha w exp(-w/w_c) with alpha = 0.5,
w_c = 5, T = 0.2; dt = 0.05, dkmax = 200, epsrel = 1e-7, 1000 steps,
complex128. A "step" is one PT-TEMPO ``compute_step`` (zip-up of the
201-node influence MPO into the device-resident MPS plus the truncating
sweep). The influence tensors are computed on the host before the timed
region; ``--warmup`` steps run untimed, then ``--steps`` steps are timed
between device synchronisations. The bond dimensions of this workload grow
by about 11 per step during the first ~200 steps, so the timed window is the
early part of the run; the JSON line names it.

The CPU baseline times the FP64 oracle (numpy/LAPACK) on the same window of
the same workload on the host cores.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ALPHA, CUTOFF, TEMP, DT, DKMAX, EPSREL, NSTEPS = 0.5, 5.0, 0.2, 0.05, 200, \
    1e-7, 1000
FP64_MFMA_PEAK_TFLOPS = 78.6


def build_model():
    import numpy as np
    from oqupy_amd import operators as op
    from oqupy_amd.bath import Bath
    from oqupy_amd.bath_correlations import PowerLawSD
    from oqupy_amd.tempo import TempoParameters
    corr = PowerLawSD(alpha=ALPHA, zeta=1.0, cutoff=CUTOFF,
                      cutoff_type="exponential", temperature=TEMP)
    bath = Bath(0.5 * op.sigma("z"), corr)
    params = TempoParameters(dt=DT, dkmax=DKMAX, epsrel=EPSREL)
    return bath, params, np


def influence_table(bath, params):
    from oqupy_amd.tempo import influence_matrix
    corr = bath.correlations
    return {dk: influence_matrix(dk, params, corr, bath.coupling_acomm,
                                 bath.coupling_comm)
            for dk in range(DKMAX + 1)}


def gpu_run(warmup, steps, table, precision="c128"):
    import torch
    import oqupy_amd as oqupy
    bath, params, np = build_model()
    ptt = oqupy.PtTempo(bath, 0.0, NSTEPS * DT + 1e-9, params,
                        backend_config={"backend": "mi355x",
                                        "precision": precision,
                                        "chi_cap": 8192})
    assert ptt.num_steps == NSTEPS
    ptt._influence = lambda dk: table.get(dk)
    backend = ptt.backend
    backend._influence = ptt._influence
    backend.initialize()
    for _ in range(warmup):
        backend.compute_step()
    torch.cuda.synchronize()
    before = backend.svd_kernel_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        backend.compute_step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    after = backend.svd_kernel_stats()
    window = backend.step_stats[-steps:]
    svd = {k: after[k] - before[k] for k in after}
    return elapsed, svd, window, backend.step


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


def cpu_baseline(warmup, steps, table, budget_s):
    """The oracle on the same window (capped at ``budget_s`` seconds of timed
    work): steps/s over the steps it completed."""
    from oracle.pt_tempo import PtTempoOracle
    import numpy as np
    pt = PtTempoOracle(2, lambda dk: table.get(dk), np.ones(4), NSTEPS, DKMAX,
                       EPSREL)
    pt.initialize()
    for _ in range(warmup):
        pt.compute_step()
    t0 = time.perf_counter()
    done = 0
    for _ in range(steps):
        pt.compute_step()
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return done / (time.perf_counter() - t0), done


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=25.0,
                    help="seconds of timed oracle work for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=("c128", "c64"), default="c128",
                    help="c64: contractions in complex64 on the matrix cores "
                         "(FP32 accumulation), SVD and cut in FP64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; a "
                          "single PT-TEMPO instance is a sequential chain of "
                          "SVDs and is not sharded across GPUs"}))
        sys.exit(2)

    bath, params, np = build_model()
    t0 = time.perf_counter()
    table = influence_table(bath, params)
    precompute_s = time.perf_counter() - t0

    elapsed, svd, window, last_step = gpu_run(args.warmup, args.steps, table,
                                              args.precision)
    value = args.steps / elapsed
    svd_avg_ms = svd["ms"] / max(svd["count"], 1)
    achieved = svd["flops"] / max(svd["ms"], 1e-12) / 1e9   # TFLOP/s
    roofline = {
        "kernel": "jacobi_kernel (blocked one-sided Jacobi SVD, FP64)",
        "bound": "mfma", "achieved": round(achieved, 4),
        "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 6),
        "traffic": None,
        "launches": svd["count"], "avg_launch_ms": round(svd_avg_ms, 4),
        "algorithmic_flops_per_launch": svd["flops"] / max(svd["count"], 1),
    }
    cpu = None
    if not args.no_cpu:
        cores = blas_threads()
        cpu_value, done = cpu_baseline(args.warmup, args.steps, table,
                                       args.cpu_budget)
        cpu = {"value": round(cpu_value, 6), "unit": "time-steps/s",
               "cores": cores, "kind": "port",
               "sample": f"FP64 oracle (numpy/LAPACK, OpenBLAS threads = "
                         f"{cores}) on steps {args.warmup + 2}.."
                         f"{args.warmup + 1 + done} of the same workload"}
    out = {
        "metric": "PT-TEMPO time-steps/s at dkmax=200",
        "value": round(value, 6), "unit": "time-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {
            "workload": (f"spin-boson PT-TEMPO alpha={ALPHA} wc={CUTOFF} "
                         f"T={TEMP} dt={DT} dkmax={DKMAX} epsrel={EPSREL} "
                         f"N={NSTEPS}; timed steps {args.warmup + 2}.."
                         f"{args.warmup + 1 + args.steps}"),
            "precision": ("plain complex128" if args.precision == "c128" else
                          "complex64 contractions (FP32 accumulation), "
                          "FP64 SVD and cut"),
            "max_bond_in_window": max(w[2] for w in window),
            "svds_per_step": round(sum(w[0] for w in window) / len(window), 1),
            "jacobi_sweeps_per_svd": round(
                sum(w[1] for w in window) / max(1, sum(w[0] for w in window)),
                2),
            "influence_precompute_s": round(precompute_s, 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
