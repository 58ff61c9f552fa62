"""PT-TEMPO backend running on the MI355X engine.

Keeps the interface of the reference's ``PtTempoBackend``
(oqupy/backends/pt_tempo_backend.py:28-316): ``initialize``, ``compute_step``,
``step``, ``num_steps``, ``get_mpo_tensor`` and ``update_process_tensor``.
The influence callable is evaluated here, on the host, and its arrays are
handed to the engine: for dk = 0 .. min(N, dkmax+1)-1 before initialisation,
and influence(-step) at every growth step (a rectangle influence when a
correlation time is added, else ``None``), exactly where the reference calls
it. The tensor chain stays in device memory.
"""

import ctypes
import weakref
from typing import Callable, Dict, List, Optional

import numpy as np
from numpy import ndarray

from oqupy_amd.backends import engine as eng


class PtTempoBackend:
    """Process-tensor TEMPO on the MI355X engine."""

    def __init__(self, dimension: int, influence: Callable[[int], ndarray],
                 process_tensor, sum_north: ndarray, sum_west: ndarray,
                 num_steps: int, dkmax: int, epsrel: float, config: Dict,
                 degeneracy_maps: Optional[List[ndarray]] = None) -> None:
        self._dimension = dimension
        self._influence = influence
        # held weakly: the finished process tensor keeps this backend (its
        # device-resident engine) alive, and a strong reference back would
        # make a cycle that only the garbage collector could free
        self._process_tensor_ref = weakref.ref(process_tensor)
        self._sum_north = np.asarray(sum_north)
        self._sum_west = np.asarray(sum_west)
        self._num_steps = num_steps
        self._dkmax = dkmax
        self._epsrel = epsrel
        self._config = eng.parse_backend_config(config)
        self._degeneracy_maps = degeneracy_maps
        self._num_infl = min(num_steps, dkmax + 1)
        self._step = None
        self._handle = None
        self.step_stats = []
        self._lib = eng.load_library()

    @property
    def step(self) -> Optional[int]:
        """The current step (None before ``initialize``)."""
        return self._step

    @property
    def num_steps(self) -> int:
        """Number of steps of the process tensor."""
        return self._num_steps

    @property
    def handle(self):
        """Opaque engine handle (valid after ``initialize``)."""
        return self._handle

    def _create(self) -> None:
        eng.require_gpu()
        cfg = eng._PtConfig(dimension=self._dimension,
                            num_steps=self._num_steps, dkmax=self._dkmax,
                            epsrel=self._epsrel,
                            precision=eng.PRECISIONS[self._config["precision"]],
                            chi_cap=self._config["chi_cap"],
                            n_north=int(self._sum_north.size))
        handle = ctypes.c_void_p()
        eng.check(self._lib.pte_pt_create(ctypes.byref(cfg),
                                          ctypes.byref(handle)))
        self._handle = handle
        eng.track(self)

    def initialize(self) -> None:
        """Build the MPO and MPS and sweep them (pt_tempo_backend.py:105-190)."""
        self._create()
        for dk in range(self._num_infl):
            infl = eng.as_c128(self._influence(dk))
            rows = infl.shape[0]
            cols = infl.shape[1] if infl.ndim == 2 else 0
            eng.check(self._lib.pte_pt_set_influence(
                self._handle, dk, eng.dptr(infl), rows, cols))
        if self._degeneracy_maps is not None:
            nmap = np.ascontiguousarray(self._degeneracy_maps[0],
                                        dtype=np.int32)
            wmap = np.ascontiguousarray(self._degeneracy_maps[1],
                                        dtype=np.int32)
            eng.check(self._lib.pte_pt_set_degeneracy_maps(
                self._handle, eng.iptr(nmap), eng.iptr(wmap), int(nmap.size)))
        self._configure_rank()
        eng.check(self._lib.pte_pt_initialize(self._handle))
        self._step = 1

    def _configure_rank(self) -> None:
        """Learned truncation-rank form: trace recording and the cap of the
        pivoted QRs (the MLP's prediction, the bond's previous step count, or
        the adaptive stop at the acceptance threshold)."""
        cfg = self._config
        if cfg["svd_log"] > 0:
            self.svd_log(cfg["svd_log"])
        if cfg["rank_trace"] > 0:
            eng.check(self._lib.pte_pt_rank_trace(self._handle,
                                                  cfg["rank_trace"]))
        predictor = cfg["rank_predictor"]
        if predictor is None:
            return
        alpha, temperature = cfg["rank_bath"]
        params = None
        if predictor == "previous":
            mode = 1
        elif predictor == "adaptive":
            mode = 3
        else:
            params = np.ascontiguousarray(predictor.params, dtype=np.float32)
            if params.size != self._lib.pte_rank_param_count():
                raise ValueError("RankPredictor parameters of the wrong size.")
            mode = 2
        eng.check(self._lib.pte_pt_set_rank_mode(
            self._handle, mode, eng.fptr(params), alpha, temperature,
            cfg["rank_oversample"], cfg["rank_margin"], cfg["rank_theta"]))

    def rank_trace(self) -> ndarray:
        """Training records of the truncations so far: one row per SVD with
        the 24 features, the target log2(need)/8, need (the pivoted-QR steps
        that reached the acceptance threshold), the kept rank and min(m, n)."""
        count = ctypes.c_int()
        eng.check(self._lib.pte_pt_get_rank_trace(self._handle, None, 0,
                                                  ctypes.byref(count)))
        width = 28
        out = np.zeros((count.value, width), dtype=np.float32)
        if count.value:
            eng.check(self._lib.pte_pt_get_rank_trace(
                self._handle, eng.fptr(out), count.value, ctypes.byref(count)))
        return out

    def rank_stats(self) -> Dict[str, int]:
        """Truncations seen by the learned form, how many were capped, how
        many fell back to the exact SVD, and the pivoted-QR steps taken."""
        vals = [ctypes.c_long() for _ in range(4)]
        eng.check(self._lib.pte_pt_rank_stats(
            self._handle, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("svds", "capped", "fallbacks", "qr_steps"),
                        (v.value for v in vals)))

    def compute_step(self) -> bool:
        """One PT-TEMPO step (pt_tempo_backend.py:192-282); returns whether
        more steps follow."""
        next_step = self._step + 1
        end_phase = next_step > self._num_steps - self._num_infl + 1
        last = None
        rows = cols = 0
        if not end_phase:
            infl = self._influence(int(0 - next_step))
            if infl is not None:
                last = eng.as_c128(infl)
                rows, cols = last.shape
        more = ctypes.c_int()
        eng.check(self._lib.pte_pt_step(self._handle, eng.dptr(last), rows,
                                        cols, ctypes.byref(more)))
        self._step = next_step
        stats = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int(),
                 ctypes.c_double())
        eng.check(self._lib.pte_pt_last_step_stats(
            self._handle, *[ctypes.byref(s) for s in stats]))
        self.step_stats.append(tuple(s.value for s in stats))
        return bool(more.value)

    def svd_log(self, max_values: int) -> None:
        """Record the truncated SVDs that follow (up to ``max_values`` kept
        singular values in total; 0 turns the log off)."""
        eng.check(self._lib.pte_pt_svd_log(self._handle, int(max_values)))

    def get_svd_log(self) -> List[tuple]:
        """The recorded SVDs in order: ``(m, n, kept singular values)``, the
        shape of the matrix the reference factors and its kept spectrum."""
        ns, nv = ctypes.c_int(), ctypes.c_long()
        eng.check(self._lib.pte_pt_get_svd_log(self._handle, None, 0, None, 0,
                                               ctypes.byref(ns),
                                               ctypes.byref(nv)))
        shapes = np.zeros(3 * max(ns.value, 1), dtype=np.int32)
        values = np.zeros(max(nv.value, 1), dtype=np.float64)
        eng.check(self._lib.pte_pt_get_svd_log(
            self._handle, eng.iptr(shapes), ns.value,
            values.ctypes.data_as(eng._DP), nv.value, ctypes.byref(ns),
            ctypes.byref(nv)))
        out, off = [], 0
        for i in range(ns.value):
            m, n, k = (int(v) for v in shapes[3 * i:3 * i + 3])
            out.append((m, n, values[off:off + k].copy()))
            off += k
        return out

    def sync_step(self) -> None:
        """Read the step counter back from the engine (after batched steps)."""
        st = ctypes.c_int()
        eng.check(self._lib.pte_pt_get_step(self._handle, ctypes.byref(st)))
        self._step = st.value

    def svd_kernel_stats(self) -> dict:
        """Accumulated launches, device milliseconds, algorithmic flops and
        bytes of the Jacobi SVD kernel."""
        count = ctypes.c_long()
        vals = [ctypes.c_double() for _ in range(3)]
        eng.check(self._lib.pte_pt_svd_kernel_stats(
            self._handle, ctypes.byref(count),
            *[ctypes.byref(v) for v in vals]))
        return {"count": count.value, "ms": vals[0].value,
                "flops": vals[1].value, "bytes": vals[2].value}

    def bond_dimensions(self) -> List[int]:
        """Bond dimensions of the device-resident MPS."""
        n = ctypes.c_int()
        eng.check(self._lib.pte_pt_num_sites(self._handle, ctypes.byref(n)))
        out = np.zeros(max(n.value - 1, 0), dtype=np.int32)
        eng.check(self._lib.pte_pt_bond_dims(self._handle, eng.iptr(out),
                                             int(out.size)))
        return [int(b) for b in out]

    def get_site(self, index: int) -> ndarray:
        """MPS site ``index`` as stored on the device, ``[left, physical,
        right]`` (any step after initialisation)."""
        n = ctypes.c_int()
        eng.check(self._lib.pte_pt_num_sites(self._handle, ctypes.byref(n)))
        bonds = [1] + self.bond_dimensions() + [1]
        if not 0 <= index < n.value:
            raise IndexError("site index out of range")
        # the physical leg is d^2 (1 for a site appended this step)
        count = bonds[index] * bonds[index + 1] * self._dimension ** 2
        out = np.empty(count, dtype=np.complex128)
        shape = np.zeros(3, dtype=np.int32)
        eng.check(self._lib.pte_pt_get_site(self._handle, int(index),
                                            eng.dptr(out), count,
                                            eng.iptr(shape)))
        size = int(np.prod(shape))
        return out[:size].reshape(tuple(int(v) for v in shape)).copy()

    def get_mpo_tensor(self, step: int) -> ndarray:
        """PT-MPO tensor ``[bond_past, bond_future, d^2]`` of ``step``
        (pt_tempo_backend.py:284-307)."""
        assert self._step is not None and self._step >= self._num_steps
        assert step < self._num_steps
        bonds = self.bond_dimensions()
        left = 1 if step == 0 else bonds[step - 1]
        right = 1 if step == self._num_steps - 1 else bonds[step]
        count = left * right * self._dimension ** 2
        out = np.empty(count, dtype=np.complex128)
        shape = np.zeros(3, dtype=np.int32)
        eng.check(self._lib.pte_pt_get_mpo_tensor(
            self._handle, step, eng.dptr(out), count, eng.iptr(shape)))
        return out.reshape(tuple(int(s) for s in shape))

    def compute_caps(self) -> List[ndarray]:
        """Caps of the finished process tensor, computed on the device
        (process_tensor.py:380-406)."""
        eng.check(self._lib.pte_pt_compute_caps(self._handle))
        caps = []
        bonds = [1] + self.bond_dimensions() + [1]
        for k in range(self._num_steps + 1):
            out = np.empty(bonds[k], dtype=np.complex128)
            length = ctypes.c_int()
            eng.check(self._lib.pte_pt_get_cap(
                self._handle, k, eng.dptr(out), bonds[k],
                ctypes.byref(length)))
            caps.append(out[:length.value].copy())
        return caps

    def update_process_tensor(self) -> None:
        """Export every PT-MPO tensor and the caps into the process tensor
        (pt_tempo_backend.py:309-316)."""
        assert self._step >= self._num_steps
        process_tensor = self._process_tensor_ref()
        assert process_tensor is not None, "the process tensor was released"
        for step in reversed(range(self._num_steps)):
            process_tensor.set_mpo_tensor(step, self.get_mpo_tensor(step))
        caps = self.compute_caps()
        for k, cap in enumerate(caps):
            process_tensor.set_cap_tensor(k, cap)
        process_tensor.attach_engine(self)

    def release(self) -> None:
        """Destroy the engine and free its device memory (idempotent)."""
        handle = getattr(self, "_handle", None)
        self._handle = None
        if handle is not None and handle.value and not eng.finalizing():
            self._lib.pte_pt_destroy(handle)

    def __del__(self) -> None:
        try:
            self.release()
        except Exception:  # interpreter shutdown
            pass
