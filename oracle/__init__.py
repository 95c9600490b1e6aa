"""CPU oracle for the weighted-quadrature matrix-free operator path.

TEST INFRASTRUCTURE ONLY.  This package restates, in plain numpy/scipy, the
reference algorithms of the hot path so that the HIP implementation in
``igaops_amd`` can be checked against them.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it,
and only as the checker / the timed CPU baseline; the product never calls it.

Modules and the reference code each one follows:

* ``matrix_free``  - L-mode Kronecker products, ``operators/utils/matrix_free.py:7-119``
* ``operators``    - B-spline and NURBS matrix-free operators,
  ``operators/bsplines_ops.py:290-429`` and ``operators/nurbs_ops.py:14-77,320-463``
* ``assembly``     - explicit sparse assembly used as an independent pin of the
  operators (same integrals, Kronecker matrices instead of mode products)
* ``fastdiag``     - spatial FD (``fastdiag/utils/space_precond.py:124-320``) and the
  space-time arrowhead FD (``fastdiag/utils/time_arrow.py:44-192``,
  ``fastdiag/utils/operations.py:5-93``, ``fastdiag/utils/time_template.py:46-71``)
* ``solvers``      - ``LinearSolver._GMRES`` with its stopping rules
  (``solvers/linear_solver.py:149-257,438-525``,
  ``solvers/utils/convergence_manager.py:13-327``) and GCRODR
  (``solvers/deflated/gcrodr.py:188-373``, ``solvers/deflated/helpers.py:13-265``)
* ``polynomial``   - the GMRES polynomial preconditioner
  (``solvers/polynomial/preconditioner.py:7-366``)

How each function is pinned (closed forms, reference known answers, brute force)
is listed in the module docstrings and exercised by ``tests/test_oracle_*.py``.
"""
