"""Chain operations of the reference NodeArray on explicit site arrays.

Sites of a matrix-product state are ``[left, physical, right]`` arrays; nodes
of a matrix-product operator are ``[left, in, out, right]`` arrays. Where the
reference has no dangling bond leg, a dimension-1 leg stands in for it: the
matrices handed to the SVD are then identical to the reference's, because a
length-1 axis changes no reshape.

Every function that truncates takes an optional ``spectra`` list, which
receives ``(rows, columns, singular values)`` of each SVD in order: the shape
of the matrix the reference factors and its whole spectrum.

Pinning: ``truncated_svd`` is checked against the rule restated in-tree by the
reference (``oqupy/mps_mpo.py:452-457``) and against hand-made spectra at the
cut boundary; zip-up and sweeps are pinned end-to-end by the golden states of
the reference's physics tests and by the brute-force path sum.
"""

from typing import List, Tuple

import numpy as np


def truncation_rank(s: np.ndarray, epsrel: float) -> int:
    """Number of singular values kept by tensornetwork 0.4.3's
    ``split_node_full_svd(max_truncation_err=epsrel, relative=True)``:
    the count of ``j`` with ``sqrt(sum_{i>=j} s_i^2) > epsrel * s_0``
    (same rule as ``oqupy/mps_mpo.py:452-457``)."""
    if s.size == 0:
        return 0
    tail = np.sqrt(np.cumsum(np.square(s[::-1])))
    return int(np.count_nonzero(tail > epsrel * s[0]))


def truncated_svd(mat: np.ndarray, epsrel: float, spectra: list = None
                  ) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Thin SVD by LAPACK gesdd (``numpy.linalg.svd``, as tensornetwork's
    numpy backend calls it), truncated by ``truncation_rank``; ``spectra``
    (optional) receives ``(rows, columns, all singular values)``."""
    u, s, vh = np.linalg.svd(mat, full_matrices=False)
    if spectra is not None:
        spectra.append((mat.shape[0], mat.shape[1], s.copy()))
    k = truncation_rank(s, epsrel)
    return u[:, :k], s[:k], vh[:k, :]


def contract_site(site: np.ndarray, node: np.ndarray,
                  carry: np.ndarray) -> np.ndarray:
    """``T[x, s, l, w] = sum_{p, r, v} A[l, p, r] W[w, p, s, v] C[x, r, v]``:
    one site of a zip-up toward the left (node_array.py:503-520 with
    ``direction='left'``)."""
    return np.einsum('lpr,wpsv,xrv->xslw', site, node, carry, optimize=True)


def contract_site_right(site: np.ndarray, node: np.ndarray,
                        carry: np.ndarray) -> np.ndarray:
    """``T[x, s, r, e] = sum_{p, l, w} C[x, l, w] A[l, p, r] W[w, p, s, e]``:
    one site of a zip-up toward the right (node_array.py:503-520 with
    ``direction='right'``)."""
    return np.einsum('xlw,lpr,wpse->xsre', carry, site, node, optimize=True)


def identity_carry(site: np.ndarray, node: np.ndarray, right: bool
                   ) -> np.ndarray:
    """The carry at the first site of a zip-up: identity over the two
    outer bond legs, fused into one leg ``x``."""
    if right:
        dl, dw = site.shape[0], node.shape[0]
    else:
        dl, dw = site.shape[2], node.shape[3]
    return np.eye(dl * dw, dtype=complex).reshape(dl * dw, dl, dw)


def zip_up_left(mps: List[np.ndarray], mpo: List[np.ndarray], epsrel: float,
                log: list = None, spectra: list = None) -> None:
    """Zip ``mpo`` into the last ``len(mpo)`` sites of ``mps``, right to left,
    with a truncated SVD at every site but the leftmost
    (NodeArray.zip_up, node_array.py:413-554, ``direction='left'``,
    ``right_index=-1``). Modifies ``mps`` in place."""
    n = len(mpo)
    first = len(mps) - n
    assert first >= 0, "MPO longer than the MPS window"
    carry = None
    for j in range(n - 1, -1, -1):
        ia = first + j
        site, node = mps[ia], mpo[j]
        if carry is None:
            carry = identity_carry(site, node, right=False)
        t = contract_site(site, node, carry)
        dx, ds, dl, dw = t.shape
        if j == 0:
            assert dw == 1
            mps[ia] = t.reshape(dx, ds, dl).transpose(2, 1, 0).copy()
            break
        u, s, vh = truncated_svd(t.reshape(dx * ds, dl * dw), epsrel,
                                 spectra)
        if log is not None:
            log.append(len(s))
        k = len(s)
        mps[ia] = u.reshape(dx, ds, k).transpose(2, 1, 0).copy()
        carry = (s[:, None] * vh).reshape(k, dl, dw)


def zip_up_right(mps: List[np.ndarray], mpo: List[np.ndarray], epsrel: float,
                 log: list = None, spectra: list = None) -> None:
    """Zip ``mpo`` (same length as ``mps``) into ``mps`` left to right, with a
    truncated SVD at every site but the rightmost (NodeArray.zip_up,
    ``direction='right'``, ``left_index=0``, ``right_index=-1``)."""
    n = len(mpo)
    assert len(mps) == n
    carry = None
    for j in range(n):
        site, node = mps[j], mpo[j]
        if carry is None:
            carry = identity_carry(site, node, right=True)
        t = contract_site_right(site, node, carry)
        dx, ds, dr, de = t.shape
        if j == n - 1:
            assert dr == 1
            mps[j] = t.reshape(dx, ds, de).copy()
            break
        u, s, vh = truncated_svd(t.reshape(dx * ds, dr * de), epsrel,
                                 spectra)
        if log is not None:
            log.append(len(s))
        k = len(s)
        mps[j] = u.reshape(dx, ds, k).copy()
        carry = (s[:, None] * vh).reshape(k, dr, de)


def sweep_right(mps: List[np.ndarray], start: int, stop: int, epsrel: float,
                log: list = None, spectra: list = None) -> None:
    """Truncated SVD sweep from site ``start`` to site ``stop`` (exclusive of
    the SVD at ``stop``), left to right (NodeArray.svd_sweep,
    node_array.py:251-273)."""
    for i in range(start, stop):
        dl, dp, dr = mps[i].shape
        u, s, vh = truncated_svd(mps[i].reshape(dl * dp, dr), epsrel,
                                 spectra)
        if log is not None:
            log.append(len(s))
        k = len(s)
        mps[i] = u.reshape(dl, dp, k)
        mps[i + 1] = np.einsum('kr,rqs->kqs', s[:, None] * vh, mps[i + 1])


def sweep_left(mps: List[np.ndarray], start: int, stop: int, epsrel: float,
               log: list = None, spectra: list = None) -> None:
    """Truncated SVD sweep from site ``start`` down to site ``stop``
    (exclusive of the SVD at ``stop``), right to left
    (NodeArray.svd_sweep, node_array.py:274-296)."""
    for i in range(start, stop, -1):
        dl, dp, dr = mps[i].shape
        mat = mps[i].reshape(dl, dp * dr).T
        u, s, vh = truncated_svd(mat, epsrel, spectra)
        if log is not None:
            log.append(len(s))
        k = len(s)
        mps[i] = u.T.reshape(k, dp, dr)
        mps[i - 1] = np.einsum('lqr,kr->lqk', mps[i - 1], s[:, None] * vh)
