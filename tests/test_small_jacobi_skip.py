"""CPU check of the pair skipping in the one-workgroup Jacobi
(small_jacobi in oqupy_amd/csrc/pte_kernels.hip): with the kernel's
round-robin order, which repeats every sweep, a pair is skipped when neither
of its columns has been rotated since the round one sweep earlier; that
round checked the same pair on the same columns and found it converged. The
claim is that skipping changes nothing: the sweeps run on a numpy model of
the kernel's rotations (same pairing, stopping rule and rotation formula up
to rounding) must give the same bits and the same sweep count with and
without the skip, on matrices of the kind the kernel sees (Y = R^H of a
pivoted QR of graded complex matrices, even and odd column counts). The
planted variant that looks back one round instead of one sweep (it skips
pairs it never checked on their current columns) must change the result,
so the test can fail. (The Jacobi corrects many such slips by itself: a lag
of one round less than a sweep, or of half a sweep, happens to give the same
bits on these matrices.)"""

import numpy as np
import pytest
import scipy.linalg as sla

EPS = 2.220446049250313e-16


def _pairing(re, rnd, q):
    """Pair q of round rnd among re (even) columns, as in the kernel."""
    lr = re - 1
    if q == 0:
        ci, cj = lr, rnd
    else:
        ci = (rnd + q) % lr
        cj = (rnd - q) % lr
    return (ci, cj) if ci < cj else (cj, ci)


def _rotation(app, aqq, apq):
    """The 2x2 Hermitian rotation {cs, sn, phase} (Rutishauser's formulas)."""
    mag = abs(apq)
    zeta = (aqq - app) / (2.0 * mag)
    t = 1.0 if zeta == 0.0 else np.sign(zeta) / (abs(zeta) + np.hypot(1.0, zeta))
    cs = 1.0 / np.hypot(1.0, t)
    return cs, cs * t, apq / mag


def _sweeps(y, skip_lag, max_sweeps=40):
    """One-sided Jacobi on the columns of y in the kernel's order; skip_lag
    is the number of rounds after which an unrotated, converged pair is
    skipped (None: never). Returns (y, sweeps)."""
    y = y.copy()
    n, r = y.shape
    tol2 = (max(n, 16) * EPS) ** 2
    abs_tol2 = (EPS * np.linalg.norm(y)) ** 2
    re = r + (r & 1)
    lr = max(re - 1, 1)
    last_rot = np.full(r, -(1 << 30))
    for sweep in range(max_sweeps):
        any_rot = False
        for rnd in range(lr):
            gr = sweep * lr + rnd
            for q in range(re // 2):
                ci, cj = _pairing(re, rnd, q)
                if cj >= r:
                    continue
                if (skip_lag is not None and gr >= skip_lag
                        and last_rot[ci] < gr - skip_lag
                        and last_rot[cj] < gr - skip_lag):
                    continue
                app = np.vdot(y[:, ci], y[:, ci]).real
                aqq = np.vdot(y[:, cj], y[:, cj]).real
                apq = np.vdot(y[:, ci], y[:, cj])
                mag2 = abs(apq) ** 2
                if not (mag2 > tol2 * abs(app * aqq)
                        and mag2 > abs_tol2 * max(app, aqq, 0.0) and mag2 > 0.0):
                    continue
                cs, sn, ph = _rotation(app, aqq, apq)
                xp = y[:, ci].copy()
                xq = y[:, cj] * np.conj(ph)
                y[:, ci] = cs * xp - sn * xq
                y[:, cj] = sn * xp + cs * xq
                last_rot[ci] = last_rot[cj] = gr
                any_rot = True
        if not any_rot:
            return y, sweep + 1
    raise AssertionError("no convergence")


def _r_factor_h(m, n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
    u, _, vh = np.linalg.svd(a, full_matrices=False)
    s = 10.0 ** (-np.linspace(0, 14, min(m, n)))
    mat = (u * s) @ vh
    _, rfac, _ = sla.qr(mat, mode="economic", pivoting=True)
    rows = np.linalg.norm(np.triu(rfac), axis=1)
    tail = np.sqrt(np.cumsum((rows ** 2)[::-1])[::-1])
    r = int(np.sum(tail > EPS * np.linalg.norm(mat)))
    return rfac[:r].conj().T.copy()


@pytest.mark.parametrize("m,n,seed", [(40, 30, 1), (30, 23, 2), (48, 48, 3)])
def test_skip_changes_nothing(m, n, seed):
    y = _r_factor_h(m, n, seed)
    re = y.shape[1] + (y.shape[1] & 1)
    lr = max(re - 1, 1)
    plain, sw_plain = _sweeps(y, None)
    skipped, sw_skip = _sweeps(y, lr)
    assert sw_plain == sw_skip
    assert np.array_equal(plain, skipped)


@pytest.mark.parametrize("m,n,seed", [(40, 30, 1), (48, 48, 3)])
def test_skip_with_a_lag_of_one_round_is_caught(m, n, seed):
    """Planted slip: comparing with the previous round instead of the same
    round one sweep earlier skips pairs that were never checked on their
    current columns, and the result moves."""
    y = _r_factor_h(m, n, seed)
    plain, _ = _sweeps(y, None)
    wrong, _ = _sweeps(y, 1)
    assert not np.array_equal(plain, wrong)
