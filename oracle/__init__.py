"""Oracle: CPU restatement of the reference SIMP hot path (TEST INFRASTRUCTURE).

This package is the correctness checker for ``toptimization_amd``. Only the
test suite, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import or run it; the product never does.

* ``oracle.simp_np`` -- FP64 NumPy restatement of every function on the
  path, following the reference files line by line (citations in each
  docstring). Small grids only (pure element loops are vectorised, but the
  PCG loop is Python).
* ``oracle/csrc/oracle.c`` -> ``oracle/liboracle.so`` -- the same algorithm
  in C with OpenMP, FP64 and FP32 builds, used for the large grids and as
  the timed CPU baseline ("reference algorithm, C restatement").

The reference itself (Taichi 1.7.4, Python 3.13) cannot run in this
environment, so the oracle is pinned instead by: the closed-form Q4 Ke
(Sigmund 2001), the reference's own unit tests (mesh / problem / material),
SciPy sparse assembly and ``spsolve``, brute-force neighbour tables, dense
filter matrices, finite differences of the compliance, brute-force replay
of the OC bisection, and cross-checks between the NumPy and C restatements.
"""
