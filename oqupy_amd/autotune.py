"""Launch-shape autotuner of the engine's SVD kernels.

The engine reads its launch shapes when an engine is created, from the
environment: ``PTE_QR_BLOCKS`` (workgroups of the pivoted QR),
``PTE_JACOBI_SLICE_ROWS`` (minimum rows of X per workgroup of a Jacobi block
pair) and ``PTE_JACOBI_INNER_SWEEPS`` (sweeps of each pair's 32 x 32
eigenproblem). ``tune`` times a PT-TEMPO window for every combination of a
grid of these (the shapes change the schedule and the rounding order of the
same arithmetic; a setting counts only if its bond dimensions equal those of
the default setting), and ``apply`` sets the best combination for the
engines created afterwards.

    python -m oqupy_amd.autotune --steps 4 --out tuned.json
"""

import argparse
import itertools
import json
import os
import time
from typing import Dict, List, Optional

import numpy as np

KNOBS = ("PTE_QR_BLOCKS", "PTE_JACOBI_SLICE_ROWS", "PTE_JACOBI_INNER_SWEEPS")
DEFAULT_GRID = {
    "PTE_QR_BLOCKS": [64, 128, 256],
    "PTE_JACOBI_SLICE_ROWS": [32, 64, 128],
    "PTE_JACOBI_INNER_SWEEPS": [1, 2],
}


def combinations(grid: Optional[Dict[str, List[int]]] = None):
    """Every assignment of the grid's values, in a fixed order."""
    grid = DEFAULT_GRID if grid is None else grid
    for key in grid:
        if key not in KNOBS:
            raise ValueError(f"unknown tuning knob {key}")
    keys = list(grid)
    for values in itertools.product(*[grid[k] for k in keys]):
        yield dict(zip(keys, values))


def apply(setting: Dict[str, int]) -> None:
    """Use `setting` for every engine created from now on."""
    for key, value in setting.items():
        if key not in KNOBS:
            raise ValueError(f"unknown tuning knob {key}")
        os.environ[key] = str(int(value))


def best(results: List[dict]) -> dict:
    """The fastest setting whose states agree with the default's."""
    ok = [r for r in results if r["agrees"]]
    if not ok:
        raise RuntimeError("no setting reproduced the default result")
    return min(ok, key=lambda r: r["seconds"])["setting"]


def _window(bath, params, n_total, warmup, steps, table):
    import torch
    import oqupy_amd as oqupy
    ptt = oqupy.PtTempo(bath, 0.0, n_total * params.dt + 1e-9, params,
                        backend_config={})
    backend = ptt.backend
    backend._influence = lambda dk: table.get(dk)
    backend.initialize()
    for _ in range(warmup):
        backend.compute_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        backend.compute_step()
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    return seconds, list(backend.bond_dimensions())


def tune(bath, params, n_total: int, warmup: int = 3, steps: int = 4,
         grid: Optional[Dict[str, List[int]]] = None, log=print) -> List[dict]:
    """Time `steps` PT-TEMPO steps (after `warmup`) of the given workload for
    every setting of the grid; each result records whether its bond
    dimensions equal those of the engine's default setting."""
    from oqupy_amd.tempo import influence_matrix
    corr = bath.correlations
    table = {dk: influence_matrix(dk, params, corr, bath.coupling_acomm,
                                  bath.coupling_comm)
             for dk in range(params.dkmax + 1)}
    saved = {k: os.environ.get(k) for k in KNOBS}
    try:
        for k in KNOBS:
            os.environ.pop(k, None)
        _, ref_bonds = _window(bath, params, n_total, warmup, steps, table)
        results = []
        for setting in combinations(grid):
            apply(setting)
            seconds, bonds = _window(bath, params, n_total, warmup, steps,
                                     table)
            res = {"setting": setting, "seconds": round(seconds, 4),
                   "steps_per_s": round(steps / seconds, 4),
                   "agrees": bonds == ref_bonds}
            results.append(res)
            log(json.dumps(res))
        return results
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def main() -> None:
    ap = argparse.ArgumentParser(description="Tune the SVD launch shapes on "
                                 "the headline PT-TEMPO workload.")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--dkmax", type=int, default=200)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    from oqupy_amd import operators as op
    from oqupy_amd.bath import Bath
    from oqupy_amd.bath_correlations import PowerLawSD
    from oqupy_amd.tempo import TempoParameters
    corr = PowerLawSD(alpha=0.5, zeta=1.0, cutoff=5.0,
                      cutoff_type="exponential", temperature=0.2)
    bath = Bath(0.5 * op.sigma("z"), corr)
    params = TempoParameters(dt=0.05, dkmax=args.dkmax, epsrel=1e-7)
    results = tune(bath, params, 1000, args.warmup, args.steps,
                   log=lambda s: print(s, flush=True))
    choice = best(results)
    summary = {"workload": f"spin-boson PT-TEMPO alpha 0.5 w_c 5 T 0.2 "
                           f"dkmax {args.dkmax} epsrel 1e-7, steps "
                           f"{args.warmup + 2}..{args.warmup + 1 + args.steps}",
               "best": choice, "results": results}
    print(json.dumps({"best": choice}))
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(summary, fh, indent=1)


if __name__ == "__main__":
    main()
