"""CPU checks of the host logic of compute_correlations(_nt): time parsing and
which combinations of times give a run (oqupy/system_dynamics.py:904-1130)."""

import numpy as np
import pytest

from oqupy_amd.correlations import _parse_times, _runs


def test_parse_times():
    assert list(_parse_times(3, 10, 0.1, 0.0)) == [3]
    assert list(_parse_times(slice(2, 5), 10, 0.1, 0.0)) == [2, 3, 4]
    assert list(_parse_times([0, 4, 10], 10, 0.1, 0.0)) == [0, 4, 10]
    assert list(_parse_times(0.42, 10, 0.1, 0.0)) == [4]
    assert list(_parse_times(0.55, 10, 0.1, 0.25)) == [3]
    assert list(_parse_times((0.2, 0.5), 10, 0.1, 0.0)) == [2, 3, 4, 5]
    assert list(_parse_times((0.5, 0.2), 10, 0.1, 0.0)) == [5, 4, 3, 2]
    with pytest.raises(IndexError):
        _parse_times(11, 10, 0.1, 0.0)
    with pytest.raises(IndexError):
        _parse_times(-0.3, 10, 0.1, 0.0)
    with pytest.raises(IndexError):
        _parse_times([12], 10, 0.1, 0.0)
    with pytest.raises(IndexError):
        _parse_times((0.0, 2.0), 10, 0.1, 0.0)
    with pytest.raises(AssertionError):
        _parse_times((0, 1), 10, 0.1, 0.0)
    with pytest.raises(TypeError):
        _parse_times("1", 10, 0.1, 0.0)


def test_runs_keep_time_ordered_combinations():
    a = np.array([0, 3, 6])
    b = np.array([2, 3, 7])
    runs = list(_runs([a, b]))
    # t_a = 0: all of b; t_a = 3: b >= 3; t_a = 6: b = 7 only
    assert [r[1] for r in runs] == [[0], [3], [6]]
    assert [list(r[2]) for r in runs] == [[2, 3, 7], [3, 7], [7]]
    assert [list(r[3]) for r in runs] == [[0, 1, 2], [1, 2], [2]]
    # three operators: (t1, t2) must be non-decreasing
    runs = list(_runs([np.array([4, 1]), np.array([2]), np.array([5])]))
    assert [r[1] for r in runs] == [[1, 2]]
    assert [r[0] for r in runs] == [(1, 0)]
    # nothing late enough
    assert list(_runs([np.array([8]), np.array([3])])) == []
