"""The allocation-free SVD entry point against the allocating one (needs a
GPU): pte_split_svd_noalloc through the C ABI on two calls in a row, the
second with a smaller matrix (the shared device scratch keeps the first
call's contents), bit-identical to pte_split_svd on the same inputs (same
kernels, same launch shapes); a matrix above its limit is refused."""

import sys

import numpy as np
import pytest

from oqupy_amd.backends import engine

pytestmark = pytest.mark.gpu


def _graded(m, n, seed, decades=13.3, eps=1e-7):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
    u, _, vh = np.linalg.svd(a, full_matrices=False)
    for _ in range(50):
        s = 10.0 ** (-np.linspace(0, decades, min(m, n)))
        tails = np.sqrt(np.cumsum(s[::-1] ** 2))[::-1]
        if np.all(np.abs(tails / (eps * s[0]) - 1) > 1e-4):
            break
        decades *= 1.01
    return (u * s) @ vh


@pytest.mark.parametrize("shapes", [((256, 200), (120, 90)),
                                    ((60, 50), (30, 20)),
                                    ((200, 256), (256, 64))])
def test_noalloc_matches_allocating_entry_point(shapes):
    for i, (m, n) in enumerate(shapes):
        print(f"[gpu-case] split_svd_noalloc {m}x{n} (call {i + 1})", flush=True)
        sys.stdout.flush()
        mat = _graded(m, n, seed=1000 * m + n)
        u0, s0, v0, w0 = engine.split_svd(mat, 1e-7)
        u1, s1, v1, w1 = engine.split_svd(mat, 1e-7, allocate=False)
        assert len(s1) == len(s0)
        assert w1 == w0
        assert np.array_equal(s1, s0)
        assert np.array_equal(u1, u0)
        assert np.array_equal(v1, v0)


def test_noalloc_refuses_oversize():
    print("[gpu-case] split_svd_noalloc oversize", flush=True)
    with pytest.raises(RuntimeError):
        engine.split_svd(np.ones((257, 8)), 1e-7, allocate=False)
