"""ctypes binding of ``libptengine.so``, the MI355X process-tensor engine.

The library is built in-tree (``python -m oqupy_amd.build``) next to this
package; loading fails loudly when it is missing or when no GPU is visible --
there is no CPU fallback. Errors returned by the C ABI are mapped to the
exception types the reference raises for the same misuse: ``ValueError`` for
bad arguments, ``AssertionError`` for calls out of order, ``RuntimeError``
for device, capacity and numerical failures.
"""

import atexit
import ctypes
import os
import sys
import weakref
from typing import Optional

import numpy as np

_LIB_NAME = "libptengine.so"
_LIB = None

# Backends that hold a live engine handle. Their engines are destroyed by an
# exit hook that runs before the interpreter finalises and before the HIP
# runtime (and any profiler attached to it) is torn down; a handle released
# any later would call hipFree / hipStreamDestroy on a finalised runtime.
_LIVE = weakref.WeakSet()

PTE_OK = 0
PTE_ERR_ARG = -1
PTE_ERR_STATE = -2
PTE_ERR_DEVICE = -3
PTE_ERR_CAPACITY = -4
PTE_ERR_NUMERIC = -5


class EngineError(RuntimeError):
    """A device, capacity or numerical failure reported by the engine."""

    def __init__(self, code: int, message: str) -> None:
        super().__init__(f"ptengine error {code}: {message}")
        self.code = code


class _PtConfig(ctypes.Structure):
    _fields_ = [("dimension", ctypes.c_int), ("num_steps", ctypes.c_int),
                ("dkmax", ctypes.c_int), ("epsrel", ctypes.c_double),
                ("precision", ctypes.c_int), ("chi_cap", ctypes.c_int),
                ("n_north", ctypes.c_int)]


class _TempoConfig(ctypes.Structure):
    _fields_ = [("dimension", ctypes.c_int), ("dkmax", ctypes.c_int),
                ("epsrel", ctypes.c_double), ("precision", ctypes.c_int),
                ("chi_cap", ctypes.c_int), ("n_north", ctypes.c_int),
                ("n_west", ctypes.c_int)]


def library_path() -> str:
    """Absolute path of the in-tree engine library."""
    return os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), _LIB_NAME)


_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)

_SIGNATURES = {
    "pte_last_error": (ctypes.c_char_p, []),
    "pte_version": (ctypes.c_char_p, []),
    "pte_pt_create": (ctypes.c_int, [ctypes.POINTER(_PtConfig),
                                     ctypes.POINTER(_P)]),
    "pte_pt_set_influence": (ctypes.c_int, [_P, ctypes.c_int, _DP,
                                            ctypes.c_int, ctypes.c_int]),
    "pte_pt_set_degeneracy_maps": (ctypes.c_int, [_P, _IP, _IP,
                                                  ctypes.c_int]),
    "pte_pt_initialize": (ctypes.c_int, [_P]),
    "pte_pt_step": (ctypes.c_int, [_P, _DP, ctypes.c_int, ctypes.c_int, _IP]),
    "pte_pt_get_step": (ctypes.c_int, [_P, _IP]),
    "pte_pt_num_sites": (ctypes.c_int, [_P, _IP]),
    "pte_pt_bond_dims": (ctypes.c_int, [_P, _IP, ctypes.c_int]),
    "pte_pt_get_mpo_tensor": (ctypes.c_int, [_P, ctypes.c_int, _DP,
                                             ctypes.c_long, _IP]),
    "pte_pt_get_site": (ctypes.c_int, [_P, ctypes.c_int, _DP, ctypes.c_long, _IP]),
    "pte_pt_compute_caps": (ctypes.c_int, [_P]),
    "pte_pt_get_cap": (ctypes.c_int, [_P, ctypes.c_int, _DP, ctypes.c_int,
                                      _IP]),
    "pte_pt_last_step_stats": (ctypes.c_int, [_P, _IP, _IP, _IP, _DP]),
    "pte_pt_svd_kernel_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_long),
                                               _DP, _DP, _DP]),
    "pte_pt_svd_log": (ctypes.c_int, [_P, ctypes.c_long]),
    "pte_pt_get_svd_log": (ctypes.c_int, [_P, _IP, ctypes.c_int, _DP, ctypes.c_long,
                                          _IP, ctypes.POINTER(ctypes.c_long)]),
    "pte_pt_destroy": (None, [_P]),
    "pte_tempo_create": (ctypes.c_int, [ctypes.POINTER(_TempoConfig),
                                        ctypes.POINTER(_P)]),
    "pte_tempo_set_influence": (ctypes.c_int, [_P, ctypes.c_int, _DP,
                                               ctypes.c_int, ctypes.c_int]),
    "pte_tempo_set_degeneracy_maps": (ctypes.c_int, [_P, _IP, _IP,
                                                     ctypes.c_int]),
    "pte_tempo_initialize": (ctypes.c_int, [_P, _DP, _DP]),
    "pte_tempo_step": (ctypes.c_int, [_P, _DP, _DP, _DP, ctypes.c_int,
                                      ctypes.c_int, _DP]),
    "pte_tempo_get_step": (ctypes.c_int, [_P, _IP]),
    "pte_tempo_bond_dims": (ctypes.c_int, [_P, _IP, ctypes.c_int, _IP]),
    "pte_tempo_destroy": (None, [_P]),
    "pte_dynamics_run": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _IP, _DP,
                                        _DP, _DP, _DP, _DP, _DP, _DP]),
    "pte_dynamics_run_engine": (ctypes.c_int, [_P, _DP, _DP, _DP, _DP, _DP]),
    "pte_split_svd": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _DP,
                                     ctypes.c_double, ctypes.c_int, _DP, _DP,
                                     _DP, _IP, _IP]),
    "pte_split_svd_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _DP,
                                             ctypes.c_double, ctypes.c_int, _DP,
                                             _DP, _DP, _IP, _IP]),
    "pte_contract": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _DP, _DP, _DP]),
    "pte_pt_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int,
                                           ctypes.POINTER(_P)]),
    "pte_pt_batch_run": (ctypes.c_int, [_P, ctypes.c_int, _IP]),
    "pte_pt_batch_stats": (ctypes.c_int, [_P] + [ctypes.POINTER(ctypes.c_long)] * 3
                           + [_DP] * 3),
    "pte_pt_batch_destroy": (None, [_P]),
    "pte_dynamics_run_engine_batch": (ctypes.c_int, [_P, ctypes.c_int, _DP, _DP,
                                                     _DP, _DP, _DP]),
    "pte_dynamics_run_engines": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, _DP,
                                                ctypes.POINTER(_DP),
                                                ctypes.POINTER(_DP), _DP, _DP]),
    "pte_pt_set_rank_mode": (ctypes.c_int, [_P, ctypes.c_int, _FP, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double]),
    "pte_pt_rank_trace": (ctypes.c_int, [_P, ctypes.c_int]),
    "pte_pt_get_rank_trace": (ctypes.c_int, [_P, _FP, ctypes.c_int, _IP]),
    "pte_pt_rank_stats": (ctypes.c_int, [_P] + [ctypes.POINTER(ctypes.c_long)] * 4),
    "pte_rank_param_count": (ctypes.c_int, []),
    "pte_rank_feature_count": (ctypes.c_int, []),
    "pte_rank_mlp_predict": (ctypes.c_int, [_FP, _FP, ctypes.c_int, _FP]),
    "pte_rank_mlp_grad": (ctypes.c_int, [_FP, _FP, _FP, ctypes.c_int, _FP, _FP]),
    "pte_rank_mlp_train": (ctypes.c_int, [_FP, _FP, _FP, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_float, ctypes.c_uint,
                                          _FP]),
    "pte_eta_batch": (ctypes.c_int, [ctypes.c_int, _DP, _IP, _DP, _DP,
                                     ctypes.c_int, _IP, _DP, _DP]),
    "pte_dynamics_run_multi": (ctypes.c_int, [ctypes.c_int] * 3 + [_IP, _DP, _DP, _DP,
                                              ctypes.POINTER(_DP),
                                              ctypes.POINTER(_DP), _DP, _DP]),
}

PRECISIONS = {"c128": 0, "c64": 1}


def exported_symbols():
    """Names of every C entry point this binding expects."""
    return sorted(_SIGNATURES)


def load_library(path: Optional[str] = None):
    """Load (once) and return the engine library with typed signatures."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    lib_path = path or library_path()
    if not os.path.exists(lib_path):
        raise ImportError(
            f"The MI355X engine library {lib_path} is not built; run "
            "`python -m oqupy_amd.build` (hipcc, gfx950).")
    # PyTorch-ROCm ships its own HIP runtime; loading it first makes the
    # engine bind to that same runtime instead of starting a second one.
    import torch  # noqa: F401
    lib = ctypes.CDLL(lib_path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
        # registered after torch's own exit hooks, so it runs before them
        atexit.register(release_all)
    return lib


def track(backend) -> None:
    """Record a backend that now owns an engine handle (see release_all)."""
    _LIVE.add(backend)


def release_all() -> None:
    """Destroy every live engine (the exit hook; also safe to call early)."""
    for backend in list(_LIVE):
        backend.release()
    _LIVE.clear()


def finalizing() -> bool:
    """True while the interpreter shuts down: engine handles still alive then
    are left to the process exit instead of calling into the runtime."""
    return sys.is_finalizing()


def check(code: int) -> None:
    """Raise the exception matching a non-zero engine return code."""
    if code == PTE_OK:
        return
    msg = load_library().pte_last_error().decode(errors="replace")
    if code == PTE_ERR_ARG:
        raise ValueError(msg)
    if code == PTE_ERR_STATE:
        raise AssertionError(msg)
    raise EngineError(code, msg)


def require_gpu() -> None:
    """Fail loudly when no HIP device is visible (no CPU fallback)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("The MI355X engine needs a visible HIP device; "
                           "none is available.")
    if not torch.cuda.is_initialized():
        torch.cuda.init()


def as_c128(array) -> np.ndarray:
    """C-contiguous complex128 copy of ``array``."""
    return np.ascontiguousarray(np.asarray(array, dtype=np.complex128))


def dptr(array: Optional[np.ndarray]):
    """``double*`` view of a complex128 (or float64) array, or NULL."""
    if array is None:
        return None
    return array.ctypes.data_as(_DP)


def iptr(array: np.ndarray):
    """``int*`` view of an int32 array."""
    return array.ctypes.data_as(_IP)


def fptr(array: Optional[np.ndarray]):
    """``float*`` view of a float32 array, or NULL."""
    if array is None:
        return None
    return array.ctypes.data_as(_FP)


def parse_backend_config(config: Optional[dict]) -> dict:
    """Validate a backend configuration dictionary; only the MI355X engine
    exists in this package."""
    from oqupy_amd.config import DEFAULT_CHI_CAP
    cfg = {} if config is None else dict(config)
    backend = cfg.get("backend", "mi355x")
    if backend != "mi355x":
        raise ValueError(f"Unknown backend '{backend}'; this package provides "
                         "only the 'mi355x' engine.")
    precision = cfg.get("precision", "c128")
    if precision not in PRECISIONS:
        raise ValueError(f"Unsupported precision '{precision}'; use 'c128' "
                         "or 'c64'.")
    chi_cap = cfg.get("chi_cap", None)
    chi_cap = DEFAULT_CHI_CAP if chi_cap is None else int(chi_cap)
    if chi_cap < 1 or chi_cap > DEFAULT_CHI_CAP:
        raise ValueError("chi_cap must be a positive C int (or None for "
                         "no limit).")
    out = {"backend": backend, "precision": precision, "chi_cap": chi_cap}
    # learned truncation-rank form (PT-TEMPO only)
    predictor = cfg.get("rank_predictor", None)
    if predictor is not None and predictor not in ("previous", "adaptive") \
            and getattr(predictor, "params", None) is None:
        raise ValueError("rank_predictor must be None, 'previous', "
                         "'adaptive' or a RankPredictor.")
    if predictor is not None and precision != "c128":
        raise ValueError("The learned rank form runs with precision 'c128'.")
    trace = int(cfg.get("rank_trace", 0))
    if trace < 0:
        raise ValueError("rank_trace must be >= 0.")
    oversample = int(cfg.get("rank_oversample", 8))
    margin = float(cfg.get("rank_margin", 0.1))
    theta = float(cfg.get("rank_theta", 0.02))
    if oversample < 0 or margin < 0.0 or not 0.0 < theta <= 1.0:
        raise ValueError("rank_oversample >= 0, rank_margin >= 0 and "
                         "0 < rank_theta <= 1 are required.")
    bath = tuple(float(v) for v in cfg.get("rank_bath", (0.0, 0.0)))
    svd_log = int(cfg.get("svd_log", 0))
    if svd_log < 0:
        raise ValueError("svd_log must be >= 0.")
    out["svd_log"] = svd_log
    influence = cfg.get("influence", "host")
    if influence not in ("host", "device"):
        raise ValueError("influence must be 'host' or 'device'.")
    out["influence"] = influence
    out.update(rank_predictor=predictor, rank_trace=trace,
               rank_oversample=oversample, rank_margin=margin,
               rank_theta=theta, rank_bath=bath)
    return out


def contract(a: np.ndarray, b: np.ndarray, precision: str = "c128"):
    """``a @ b`` by the engine's contraction kernels of the given precision."""
    require_gpu()
    lib = load_library()
    a = as_c128(a)
    b = as_c128(b)
    m, k = a.shape
    k2, n = b.shape
    if k != k2:
        raise ValueError("inner dimensions differ")
    c = np.empty((m, n), dtype=np.complex128)
    check(lib.pte_contract(PRECISIONS[precision], m, n, k, dptr(a), dptr(b),
                           dptr(c)))
    return c


def split_svd(matrix: np.ndarray, epsrel: float, chi_cap: int = 8192,
              allocate: bool = True):
    """Truncated SVD ``matrix = U diag(S) Vh`` by the engine's kernels, with
    the relative truncation rule of ``split_node_full_svd``
    (oqupy/backends/node_array.py:262). Returns ``(U, S, SVh, sweeps)`` with
    ``SVh = diag(S) Vh``. ``allocate=False`` calls the allocation-free entry
    point (matrices up to 256 x 256, workspace in a static device array)."""
    require_gpu()
    lib = load_library()
    entry = lib.pte_split_svd if allocate else lib.pte_split_svd_noalloc
    mat = as_c128(matrix)
    m, n = mat.shape
    r = min(m, n)
    u = np.empty(m * r, dtype=np.complex128)
    s = np.empty(r, dtype=np.float64)
    svh = np.empty(r * n, dtype=np.complex128)
    k = ctypes.c_int()
    sweeps = ctypes.c_int()
    check(entry(m, n, dptr(mat), float(epsrel), int(chi_cap), dptr(u),
                s.ctypes.data_as(_DP), dptr(svh), ctypes.byref(k),
                ctypes.byref(sweeps)))
    kk = k.value
    return (u[:m * kk].reshape(m, kk), s[:kk].copy(),
            svh[:kk * n].reshape(kk, n), sweeps.value)
