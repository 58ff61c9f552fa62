"""Process tensors in matrix-product-operator form.

Restates ``BaseProcessTensor``, ``TrivialProcessTensor`` and
``SimpleProcessTensor`` of ``oqupy/process_tensor.py:38-430``: PT-MPO tensors
``[bond_past, bond_future, in(, out)]``, caps and the basis transforms of a
non-diagonal coupling. A ``SimpleProcessTensor`` produced by ``PtTempo`` also
keeps a reference to the engine that computed it, so that ``compute_dynamics``
contracts it where it already lives, in device memory. The HDF5 file form
(``FileProcessTensor``) is not part of this package.
"""

from abc import ABC, abstractmethod
from typing import Optional, Text, Union

import numpy as np
from numpy import ndarray

from oqupy_amd import util
from oqupy_amd.base_api import BaseAPIClass
from oqupy_amd.config import NpDtype


class BaseProcessTensor(BaseAPIClass, ABC):
    """Abstract process tensor (oqupy/process_tensor.py:38-197)."""

    def __init__(self, hilbert_space_dimension: int,
                 dt: Optional[float] = None,
                 transform_in: Optional[ndarray] = None,
                 transform_out: Optional[ndarray] = None,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        self._hs_dim = hilbert_space_dimension
        self._dt = dt
        self._rho_dim = self._hs_dim ** 2
        self._trace = (np.identity(self._hs_dim, dtype=NpDtype)
                       / np.sqrt(float(self._hs_dim))).flatten()
        self._trace_square = self._trace ** 2
        if transform_in is not None:
            tin = np.array(transform_in, dtype=NpDtype)
            assert tin.ndim == 2 and tin.shape[0] == self._rho_dim
            self._in_dim = tin.shape[1]
            self._transform_in = tin
            self._trace_in = self._trace @ tin
        else:
            self._in_dim = self._rho_dim
            self._transform_in = None
            self._trace_in = self._trace
        if transform_out is not None:
            tout = np.array(transform_out, dtype=NpDtype)
            assert tout.ndim == 2 and tout.shape[1] == self._rho_dim
            self._out_dim = tout.shape[0]
            self._transform_out = tout
            self._trace_out = tout @ self._trace
        else:
            self._out_dim = self._rho_dim
            self._transform_out = None
            self._trace_out = self._trace
        super().__init__(name, description)

    @abstractmethod
    def __len__(self) -> int:
        """Number of time steps."""

    @property
    def hilbert_space_dimension(self) -> int:
        """Dimension of the system Hilbert space."""
        return self._hs_dim

    @property
    def dt(self) -> Optional[float]:
        """Time step length."""
        return self._dt

    @property
    def max_step(self) -> Union[int, float]:
        """Maximal number of time steps."""
        return len(self)

    @property
    def transform_in(self) -> Optional[ndarray]:
        """System basis -> process tensor basis super-operator."""
        return self._transform_in

    @property
    def transform_out(self) -> Optional[ndarray]:
        """Process tensor basis -> system basis super-operator."""
        return self._transform_out

    def set_initial_tensor(self, initial_tensor: Optional[ndarray] = None):
        raise RuntimeError(f"Class {type(self).__name__} doesn't allow to set "
                           "the initial tensor.")

    def set_mpo_tensor(self, step: int, tensor: Optional[ndarray] = None):
        raise RuntimeError(f"Class {type(self).__name__} doesn't allow to set "
                           "the mpo tensors.")

    def set_cap_tensor(self, step: int, tensor: Optional[ndarray] = None):
        raise RuntimeError(f"Class {type(self).__name__} doesn't allow to set "
                           "the cap tensors.")

    @abstractmethod
    def get_initial_tensor(self) -> Optional[ndarray]:
        """The (possibly correlated) initial state."""

    @abstractmethod
    def get_mpo_tensor(self, step: int,
                       transformed: Optional[bool] = True) -> ndarray:
        """PT-MPO tensor of ``step`` as ``[past, future, in, out]``."""

    @abstractmethod
    def get_cap_tensor(self, step: int) -> Optional[ndarray]:
        """Cap vector terminating the PT-MPO at ``step``."""

    @abstractmethod
    def get_bond_dimensions(self) -> Optional[ndarray]:
        """Bond dimensions of the PT-MPO."""


class TrivialProcessTensor(BaseProcessTensor):
    """Process tensor of an absent environment
    (oqupy/process_tensor.py:199-247)."""

    def __init__(self, hilbert_space_dimension: Optional[int] = 1,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        super().__init__(hilbert_space_dimension=hilbert_space_dimension,
                         name=name, description=description)

    def __len__(self) -> int:
        return 0

    @property
    def max_step(self) -> Union[int, float]:
        return float('inf')

    def get_initial_tensor(self) -> None:
        return None

    def get_mpo_tensor(self, step: int, transformed: Optional[bool] = True):
        return None

    def get_cap_tensor(self, step: int) -> ndarray:
        return np.array([1.0], dtype=NpDtype)

    def get_bond_dimensions(self) -> None:
        return None


class SimpleProcessTensor(BaseProcessTensor):
    """PT-MPO held in memory (oqupy/process_tensor.py:249-430)."""

    def __init__(self, hilbert_space_dimension: int,
                 dt: Optional[float] = None,
                 transform_in: Optional[ndarray] = None,
                 transform_out: Optional[ndarray] = None,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        self._initial_tensor = None
        self._mpo_tensors = []
        self._cap_tensors = []
        self._engine = None
        super().__init__(hilbert_space_dimension, dt, transform_in,
                         transform_out, name, description)

    def __len__(self) -> int:
        return len(self._mpo_tensors)

    def attach_engine(self, backend) -> None:
        """Remember the engine that holds these tensors in device memory."""
        self._engine = backend

    @property
    def engine(self):
        """The engine holding this process tensor on the device, or None."""
        return self._engine

    def set_initial_tensor(self, initial_tensor: Optional[ndarray] = None):
        """oqupy/process_tensor.py:275-286."""
        if initial_tensor is None:
            self._initial_tensor = None
        else:
            self._initial_tensor = np.array(initial_tensor, dtype=NpDtype)

    def set_mpo_tensor(self, step: int, tensor: Optional[ndarray] = None):
        """Store a copy of the tensor of ``step``; a rank-3 tensor has an
        implicit delta between its input and output legs."""
        length = len(self._mpo_tensors)
        if step >= length:
            self._mpo_tensors.extend([None] * (step - length + 1))
        self._mpo_tensors[step] = np.array(tensor, dtype=NpDtype)
        self._engine = None

    def set_cap_tensor(self, step: int, tensor: Optional[ndarray] = None):
        """Store the cap of ``step``."""
        length = len(self._cap_tensors)
        if step >= length:
            self._cap_tensors.extend([None] * (step - length + 1))
        self._cap_tensors[step] = np.array(tensor, dtype=NpDtype)

    def get_initial_tensor(self) -> Optional[ndarray]:
        return self._initial_tensor

    def get_mpo_tensor(self, step: int,
                       transformed: Optional[bool] = True) -> ndarray:
        """Rank-4 tensor of ``step`` with the basis transforms applied when
        ``transformed`` (oqupy/process_tensor.py:308-356)."""
        if step >= len(self._mpo_tensors) or step < 0:
            raise IndexError("Process tensor index out of bound. ")
        tensor = self._mpo_tensors[step]
        if tensor.ndim == 3:
            tensor = util.create_delta(tensor, [0, 1, 2, 2])
        if transformed is False:
            return tensor
        if self._transform_in is not None:
            tensor = np.dot(np.moveaxis(tensor, -2, -1),
                            self._transform_in.T)
            tensor = np.moveaxis(tensor, -1, -2)
        if self._transform_out is not None:
            tensor = np.dot(tensor, self._transform_out)
        return tensor

    def get_raw_mpo_tensor(self, step: int) -> ndarray:
        """The tensor of ``step`` as stored (rank 3 or 4)."""
        return self._mpo_tensors[step]

    def get_cap_tensor(self, step: int) -> Optional[ndarray]:
        if step >= len(self._cap_tensors) or step < 0:
            return None
        return self._cap_tensors[step]

    def get_bond_dimensions(self) -> ndarray:
        """Bond dimensions (oqupy/process_tensor.py:367-378)."""
        dims = [0 if t is None else t.shape[0] for t in self._mpo_tensors]
        dims.append(self._mpo_tensors[-1].shape[1])
        return np.array(dims)

    def compute_caps(self) -> None:
        """All caps backward from ``[1]``:
        ``c_k[a] = sum_{b,i} T_k[a,b,i] c_{k+1}[b] tr2[i]`` for rank-3 tensors
        (with the traces of the transforms for rank-4 tensors)
        (oqupy/process_tensor.py:380-406). Tensors handed over by the engine
        already come with their caps, computed on the device."""
        caps = [np.array([1.0], dtype=NpDtype)]
        for step in reversed(range(len(self))):
            ten = self._mpo_tensors[step]
            if ten.ndim == 3:
                cap = np.einsum('abi,b,i->a', ten, caps[0],
                                self._trace_square)
            else:
                cap = np.einsum('abio,b,i,o->a', ten, caps[0],
                                self._trace_in, self._trace_out)
            caps.insert(0, cap)
        self._cap_tensors = caps
