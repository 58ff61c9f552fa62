"""``compute_dynamics`` with process tensors (FP64 oracle).

Restates ``oqupy/system_dynamics.py:41-182``: per step the pre-measurement
control is applied (``controls(step)``, lines 131-138), the state (with one
bond leg per process tensor) is recorded through the caps
(``_apply_caps``, 643-651), the post-measurement control is applied (151-154),
then the state is propagated by the first half-step super-operator
(``_apply_system_superoperator``, 631-640), by each PT-MPO tensor
(``_apply_pt_mpos``, 654-700, with the rank-3 -> rank-4 delta of
``oqupy/process_tensor.py:345-355`` and its optional basis transforms) and by
the second half-step; after the last step only the pre-measurement control
of step N is applied before the final readout.

Pinning: golden states of the reference physics tests, the exact dephasing
formula, and the two-environments-at-alpha equals one-at-2-alpha invariant.
"""

from typing import Callable, List, Optional, Sequence

import numpy as np


def pt_transforms(unitary: np.ndarray):
    """Basis transforms of a process tensor whose bath coupling is diagonal
    in the basis of ``unitary`` (columns = eigenvectors), as
    ``oqupy/pt_tempo.py:157-174`` builds them: ``transform_in`` takes a
    row-major vectorised system state rho to vec(U^H rho U) and
    ``transform_out`` takes it back (vec(U rho U^H)), both stored transposed
    for the leg contractions of ``expand_mpo``; None for a diagonal coupling.
    Row-major vec(A X B) = kron(A, B^T) vec(X)."""
    u = np.asarray(unitary, dtype=complex)
    if np.allclose(u, np.identity(u.shape[0])):
        return None, None
    to_eigen = np.kron(u.conj().T, u.T)
    from_eigen = np.kron(u, u.conj())
    return to_eigen.T, from_eigen.T


def expand_mpo(tensor: np.ndarray, transform_in: Optional[np.ndarray],
               transform_out: Optional[np.ndarray]) -> np.ndarray:
    """Rank-3 ``[a, b, i]`` PT tensor to the rank-4 ``[a, b, in, out]`` form
    with the basis transforms applied (process_tensor.py:326-356)."""
    if tensor.ndim == 3:
        a, b, n = tensor.shape
        t4 = np.zeros((a, b, n, n), dtype=complex)
        t4[:, :, np.arange(n), np.arange(n)] = tensor
    else:
        t4 = tensor
    if transform_in is not None:
        t4 = np.einsum('abio,ji->abjo', t4, transform_in)
    if transform_out is not None:
        t4 = np.einsum('abio,op->abip', t4, transform_out)
    return t4


def compute_dynamics(propagators: Callable[[int], tuple],
                     initial_state: np.ndarray, num_steps: int,
                     pt_tensors: Sequence[List[np.ndarray]],
                     pt_caps: Sequence[List[np.ndarray]],
                     transforms: Optional[Sequence[tuple]] = None,
                     controls: Optional[Callable[[int], tuple]] = None
                     ) -> List[np.ndarray]:
    """States ``rho(t_0) ... rho(t_N)`` (as d x d matrices). ``controls(step)``
    returns the (pre, post) measurement super-operators of ``step`` (None
    for none), as ``Control.get_controls`` (oqupy/control.py)."""
    d = initial_state.shape[0]
    n_env = len(pt_tensors)
    if transforms is None:
        transforms = [(None, None)] * n_env
    state = np.array(initial_state, dtype=complex).reshape(
        [1] * n_env + [d * d])
    states = []

    def readout(state, step):
        vec = state
        for env in range(n_env):
            vec = np.tensordot(pt_caps[env][step], vec, axes=([0], [0]))
        return vec.reshape(d, d)

    def apply(state, sup_op):
        if sup_op is None:
            return state
        return np.tensordot(state, np.asarray(sup_op).T, axes=([-1], [0]))

    for step in range(num_steps):
        pre, post = controls(step) if controls is not None else (None, None)
        state = apply(state, pre)
        states.append(readout(state, step))
        state = apply(state, post)
        first, second = propagators(step)
        state = np.tensordot(state, first.T, axes=([-1], [0]))
        for env in range(n_env):
            t4 = expand_mpo(pt_tensors[env][step], *transforms[env])
            # contract bond leg `env` and the state leg with [a, ., in, .]
            state = np.moveaxis(state, env, 0)
            shp = state.shape
            mat = state.reshape(shp[0], -1, shp[-1])      # [a, rest, in]
            new = np.einsum('axi,abio->xbo', mat, t4)      # [rest, b, out]
            new = new.reshape(shp[1:-1] + (t4.shape[1], t4.shape[3]))
            new = np.moveaxis(new, -2, 0)                  # [b, rest..., out]
            state = np.moveaxis(new, 0, env)
        state = np.tensordot(state, second.T, axes=([-1], [0]))
    if controls is not None:
        state = apply(state, controls(num_steps)[0])
    states.append(readout(state, num_steps))
    return states
