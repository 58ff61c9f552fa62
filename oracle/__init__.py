"""FP64 CPU oracle of the process-tensor path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy/LAPACK, the reference algorithms that
the MI355X engine replaces:

* ``oracle.chain``      -- the truncated SVD rule of tensornetwork 0.4.3's
  ``split_node_full_svd(relative=True)`` and the chain operations of
  ``oqupy/backends/node_array.py`` (zip-up, SVD sweeps) on explicit
  ``[left, physical, right]`` site arrays;
* ``oracle.pt_tempo``   -- ``PtTempoBackend`` (oqupy/backends/pt_tempo_backend.py)
  plus the process-tensor export and caps (oqupy/process_tensor.py:380-406);
* ``oracle.tempo``      -- ``BaseTempoBackend.compute_system_step``
  (oqupy/backends/tempo_backend.py:379-575);
* ``oracle.dynamics``   -- ``compute_dynamics`` with process tensors
  (oqupy/system_dynamics.py:41-182, 631-700);
* ``oracle.path_sum``   -- brute-force, untruncated contraction of the same
  discretised influence functional, for tiny inputs.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed CPU
baseline). The product package ``oqupy_amd`` never imports it.

Pinning (what makes the oracle trustworthy) is listed per function in the
module headers and tested in ``tests/test_oracle_*.py``.
"""
