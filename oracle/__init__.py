"""CPU oracle for the bucket-elimination path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product package imports this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it, and
only as the checker (or as the timed CPU baseline), never as a code path
of the engine.

Contents
--------
* ``np_framework``  -- numpy restatement of the reference numpy bucket backend
  (``process_bucket`` / ``get_sliced_buckets`` / ``get_result_data``,
  ``qtensor/ProcessingFrameworks.py:23-37``; arithmetic of the un-vendored
  ``qtree.np_framework`` as fixed by ``ProcessingFrameworks.py:45-147`` and the
  outputs in ``examples/Qtree Basic usage.ipynb``), in complex128 or complex64.
* ``statevector``   -- an independent full-statevector QAOA simulator (N <= 20)
  that pins composers, circ2buckets, ordering and elimination end to end.
* ``closed_form``   -- the p = 1 MaxCut edge expectation in closed form
  (Wang, Hadfield, Jiang, Rieffel, PRA 97, 022304 (2018), Thm. 1).

Pinning (see ``tests/test_oracle.py``): golden fixtures F1-F3 from the
reference notebooks, brute-force loops over all index assignments for single
buckets, the statevector, and the closed form.
"""
