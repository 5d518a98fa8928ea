"""CPU oracle for the entropic Sinkhorn hot path -- TEST INFRASTRUCTURE ONLY.

This package is a float64 NumPy restatement of the reference algorithms
(POT 0.9.7.dev0), written from their published behaviour and cited file:line
by file:line.  It is the checker the GPU path is compared against:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it;
* the product package ``pot_amd`` never imports it and never falls back to it.

Modules
-------
``oracle.sinkhorn``  ``sinkhorn``, ``sinkhorn2``, ``sinkhorn_knopp``, ``sinkhorn_log``
                     (ot/bregman/_sinkhorn.py:21-918)
``oracle.greenkhorn`` ``greenkhorn`` (ot/bregman/_sinkhorn.py:921-1102)
``oracle.batch``     ``ot.batch`` linear solvers (ot/batch/_linear.py, ot/batch/_utils.py)
``oracle.utils``     ``dist``, ``euclidean_distances``, ``unif``, ``dist0``,
                     ``make_1D_gauss`` (ot/utils.py, ot/datasets.py)
``oracle.warmnet``   float64 forward/backward of the warm-start network (no
                     counterpart in the reference; pinned by finite differences)

Pinning status of every function is listed in each module header.
"""
