"""CPU checks of the autotuner's bookkeeping (no GPU)."""

import os

import pytest

from oqupy_amd import autotune


def test_combinations_and_apply(monkeypatch):
    combos = list(autotune.combinations())
    assert len(combos) == 3 * 3 * 2
    assert combos[0] == {"PTE_QR_BLOCKS": 64, "PTE_JACOBI_SLICE_ROWS": 32,
                         "PTE_JACOBI_INNER_SWEEPS": 1}
    with pytest.raises(ValueError):
        list(autotune.combinations({"PTE_UNKNOWN": [1]}))
    monkeypatch.delenv("PTE_QR_BLOCKS", raising=False)
    autotune.apply({"PTE_QR_BLOCKS": 128})
    assert os.environ["PTE_QR_BLOCKS"] == "128"
    monkeypatch.delenv("PTE_QR_BLOCKS")


def test_best_prefers_agreeing_fastest():
    res = [{"setting": {"a": 1}, "seconds": 1.0, "agrees": True},
           {"setting": {"a": 2}, "seconds": 0.5, "agrees": False},
           {"setting": {"a": 3}, "seconds": 0.8, "agrees": True}]
    assert autotune.best(res) == {"a": 3}
    with pytest.raises(RuntimeError):
        autotune.best([res[1]])
