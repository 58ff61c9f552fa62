"""Tensor-layout helpers, input checks and progress reporting.

``create_delta`` and ``add_singleton`` define every tensor layout of the
process-tensor path; they follow ``oqupy/util.py:30-84`` (same index
semantics), with the delta expansion done by one fancy-index assignment instead
of a Python loop over all entries.
"""

import sys
import time
from typing import Any, List, Optional, Text

import numpy as np
from numpy import ndarray

from oqupy_amd.config import PROGRESS_TYPE


def create_delta(tensor: ndarray, index_scrambling: List[int]) -> ndarray:
    """Expand ``tensor`` into a larger tensor with Kronecker deltas.

    The result ``R`` has ``R.shape[k] = tensor.shape[index_scrambling[k]]`` and
    ``R[a[s_0], a[s_1], ...] = tensor[a]`` for every index tuple ``a`` of
    ``tensor`` (all other entries are zero), as in ``oqupy/util.py:30-57``.
    """
    tensor = np.asarray(tensor)
    shape = tensor.shape
    out_shape = tuple(shape[i] for i in index_scrambling)
    out = np.zeros(out_shape, dtype=tensor.dtype)
    grids = np.indices(shape).reshape(len(shape), -1)
    out[tuple(grids[i] for i in index_scrambling)] = tensor.reshape(-1)
    return out


def add_singleton(tensor: ndarray, index: Optional[int] = -1,
                  copy: Optional[bool] = True) -> ndarray:
    """Insert a length-1 axis at position ``index`` (list-insert semantics,
    ``oqupy/util.py:72-84``)."""
    ten = np.array(tensor, copy=True) if copy else tensor
    shape = list(ten.shape)
    shape.insert(index, 1)
    ten.shape = tuple(shape)
    return ten


def check_convert(variable: Any, conv_type: Any, name: Text = None,
                  msg: Text = None):
    """Convert ``variable`` to ``conv_type`` or raise TypeError
    (oqupy/util.py:96-108)."""
    try:
        return conv_type(variable)
    except Exception as err:
        label = f"`{name}`" if name is not None else ""
        raise TypeError(f"Variable {label} must be type "
                        f"`{conv_type.__name__}`." + (msg or "")) from err


def check_true(expr: bool, msg: Text = None) -> None:
    """Raise ValueError unless ``expr`` (oqupy/util.py:110-115)."""
    if not expr:
        raise ValueError(msg or "")


def check_isinstance(variable: Any, types: Any, name: Text = None,
                     msg: Text = None) -> None:
    """Raise TypeError unless ``variable`` is one of ``types``
    (oqupy/util.py:117-134)."""
    types = types if isinstance(types, tuple) else (types,)
    if not isinstance(variable, types):
        label = f"`{name}`" if name is not None else ""
        names = " or ".join(f"`{t.__name__}`" for t in types)
        raise TypeError(f"Variable {label} is not of the type {names}. "
                        + (msg or ""))


class _Progress:
    """Progress reporter used as a context manager (oqupy/util.py:138-295).

    ``'silent'`` prints nothing; ``'simple'`` and ``'bar'`` print one line per
    update (no timer thread) and the elapsed time on exit.
    """

    def __init__(self, max_value: int, title: Optional[Text] = None,
                 kind: Text = "silent", file=None) -> None:
        self.max_value = max_value
        self.title = title
        self.kind = kind
        self.file = file if file is not None else sys.stdout
        self._start = None
        self._last = -1

    def enter(self) -> None:
        self._start = time.perf_counter()
        if self.kind != "silent" and self.title is not None:
            print(self.title, file=self.file)

    def exit(self) -> None:
        if self.kind != "silent":
            elapsed = time.perf_counter() - self._start
            print(f"Elapsed time: {elapsed:.1f}s", file=self.file)

    def update(self, step: Optional[int] = None) -> None:
        if self.kind == "silent" or step is None:
            return
        pct = int(100 * step / max(1, self.max_value))
        if pct != self._last:
            self._last = pct
            print(f"{pct:3d}%  {step} of {self.max_value}", file=self.file)

    def __enter__(self):
        self.enter()
        return self

    def __exit__(self, *exc) -> None:
        self.exit()


def get_progress(progress_type: Text = None):
    """Return a progress factory ``(max_value, title) -> reporter``
    (oqupy/util.py:288-295)."""
    kind = PROGRESS_TYPE if progress_type is None else progress_type
    if kind not in ("silent", "simple", "bar"):
        raise ValueError(f"Unknown progress type '{kind}'.")
    return lambda max_value, title=None: _Progress(max_value, title, kind)
