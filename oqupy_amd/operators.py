"""Spin operators, states and super-operators.

Conventions follow the reference's ``oqupy/operators.py:24-153``: density
matrices are vectorised row-major (``rho.reshape(d*d)``), so that the map
``rho -> A rho B`` is the super-operator ``kron(A, B.T)``.
"""

from typing import Text

import numpy as np
from numpy import ndarray

from oqupy_amd.config import NpDtype

_PAULI = {
    "id": ((1, 0), (0, 1)),
    "x": ((0, 1), (1, 0)),
    "y": ((0, -1j), (1j, 0)),
    "z": ((1, 0), (0, -1)),
    "+": ((0, 1), (0, 0)),
    "-": ((0, 0), (1, 0)),
}

_SPIN_STATES = {
    "up": ((1, 0), (0, 0)),
    "down": ((0, 0), (0, 1)),
    "z+": ((1, 0), (0, 0)),
    "z-": ((0, 0), (0, 1)),
    "x+": ((0.5, 0.5), (0.5, 0.5)),
    "x-": ((0.5, -0.5), (-0.5, 0.5)),
    "y+": ((0.5, -0.5j), (0.5j, 0.5)),
    "y-": ((0.5, 0.5j), (-0.5j, 0.5)),
    "mixed": ((0.5, 0.0), (0.0, 0.5)),
}


def identity(n: int) -> ndarray:
    """The n x n identity (oqupy/operators.py:40-54)."""
    return np.identity(n, dtype=NpDtype)


def sigma(name: Text) -> ndarray:
    """Pauli matrix ``'id'``, ``'x'``, ``'y'``, ``'z'``, ``'+'`` or ``'-'``
    (oqupy/operators.py:59-72)."""
    return np.array(_PAULI[name], dtype=NpDtype)


def spin_dm(name: Text) -> ndarray:
    """Spin-1/2 density matrix by name (oqupy/operators.py:75-89)."""
    return np.array(_SPIN_STATES[name], dtype=NpDtype)


def destroy(n: int) -> ndarray:
    """Bosonic annihilation operator on n levels (oqupy/operators.py:108-122)."""
    return np.diag(np.sqrt(np.arange(1, n), dtype=NpDtype), 1)


def create(n: int) -> ndarray:
    """Bosonic creation operator on n levels (oqupy/operators.py:92-105)."""
    return destroy(n).T


def commutator(operator: ndarray) -> ndarray:
    """Super-operator of ``rho -> [A, rho]`` (oqupy/operators.py:127-131)."""
    eye = np.identity(operator.shape[0])
    return np.kron(operator, eye) - np.kron(eye, operator.T)


def acommutator(operator: ndarray) -> ndarray:
    """Super-operator of ``rho -> {A, rho}`` (oqupy/operators.py:133-137)."""
    eye = np.identity(operator.shape[0])
    return np.kron(operator, eye) + np.kron(eye, operator.T)


def left_super(operator: ndarray) -> ndarray:
    """Super-operator of ``rho -> A rho`` (oqupy/operators.py:139-142)."""
    return np.kron(operator, np.identity(operator.shape[0]))


def right_super(operator: ndarray) -> ndarray:
    """Super-operator of ``rho -> rho A`` (oqupy/operators.py:144-147)."""
    return np.kron(np.identity(operator.shape[0]), operator.T)


def left_right_super(left_operator: ndarray, right_operator: ndarray) -> ndarray:
    """Super-operator of ``rho -> A rho B`` (oqupy/operators.py:149-153)."""
    return np.kron(left_operator, right_operator.T)
