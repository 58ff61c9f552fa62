"""The bath: a coupling operator plus its auto-correlations.

Restates ``oqupy/bath.py:28-164``: the coupling operator is diagonalised (if it
is not diagonal already), and the diagonals of its commutator and
anti-commutator super-operators, which enter every influence tensor, are kept,
together with the degeneracy maps used by ``unique=True``.
"""

from copy import copy
from typing import Optional, Text

import numpy as np
from numpy import ndarray

from oqupy_amd.base_api import BaseAPIClass
from oqupy_amd.bath_correlations import BaseCorrelations
from oqupy_amd.config import DEFAULT_TOLERANCE_DEGENERACY, NpDtype
from oqupy_amd.operators import acommutator, commutator


class Bath(BaseAPIClass):
    """A bath coupled to the system through a Hermitian operator
    (oqupy/bath.py:28-155)."""

    def __init__(self, coupling_operator: ndarray,
                 correlations: BaseCorrelations,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        try:
            op = np.array(coupling_operator, dtype=NpDtype)
            op.setflags(write=False)
        except Exception as err:
            raise AssertionError("Coupling operator must be numpy array") \
                from err
        assert op.ndim == 2, "Coupling operator must be a matrix."
        assert op.shape[0] == op.shape[1], \
            "Coupling operator must be a square matrix."
        assert np.allclose(op.conjugate().T, op), \
            "Coupling operator must be a hermitian matrix."
        self._dimension = op.shape[0]

        if np.allclose(np.diag(op.diagonal()), op):
            self._coupling_operator = op
            self._unitary = np.identity(self._dimension)
        else:
            vals, vecs = np.linalg.eig(op)
            self._coupling_operator = np.diag(vals)
            self._unitary = vecs
            assert np.allclose(op, vecs @ np.diag(vals) @ vecs.conjugate().T)

        self._coupling_comm = commutator(self._coupling_operator).diagonal()
        self._coupling_acomm = acommutator(self._coupling_operator).diagonal()
        self._north_degeneracy_map = _row_degeneracy(
            [self._coupling_comm, self._coupling_acomm])
        self._west_degeneracy_map = _row_degeneracy([self._coupling_comm])

        if not isinstance(correlations, BaseCorrelations):
            raise AssertionError("Correlations must be an instance of a "
                                 "subclass of BaseCorrelations.")
        self._correlations = copy(correlations)
        super().__init__(name, description)

    def __str__(self) -> Text:
        return (super().__str__()
                + f"  dimension     = {self.dimension} \n"
                + f"  correlations  = {self.correlations.name} \n")

    @property
    def coupling_operator(self) -> ndarray:
        """The diagonalised coupling operator."""
        return self._coupling_operator.copy()

    @property
    def unitary_transform(self) -> ndarray:
        """Unitary U with ``S = U diag U^dagger``."""
        return self._unitary.copy()

    @property
    def dimension(self) -> int:
        """Hilbert space dimension."""
        return copy(self._dimension)

    @property
    def correlations(self) -> BaseCorrelations:
        """The bath correlations."""
        return copy(self._correlations)

    @property
    def coupling_acomm(self) -> ndarray:
        """Diagonal of the anti-commutator super-operator of the coupling."""
        return self._coupling_acomm.copy()

    @property
    def coupling_comm(self) -> ndarray:
        """Diagonal of the commutator super-operator of the coupling."""
        return self._coupling_comm.copy()

    @property
    def north_degeneracy_map(self) -> ndarray:
        """Map to the distinct (difference, sum) pairs of coupling
        eigenvalues."""
        return copy(self._north_degeneracy_map)

    @property
    def west_degeneracy_map(self) -> ndarray:
        """Map to the distinct differences of coupling eigenvalues."""
        return copy(self._west_degeneracy_map)


def _row_degeneracy(rows) -> ndarray:
    """Index of each column of ``rows`` among the distinct columns (after
    rounding), as ``np.unique(..., return_inverse=True)`` orders them
    (oqupy/bath.py:158-164)."""
    mat = np.array(rows).round(decimals=DEFAULT_TOLERANCE_DEGENERACY)
    inverse = np.unique(mat.T, return_inverse=True, axis=0)[1]
    return np.asarray(inverse).reshape(-1)
