"""CPU restatement of the reference's solve loop -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it, and only to compute expected values (or to time the CPU
baseline).  The product (``voxscatter_amd``) never imports it and has no CPU fallback.

It restates, in NumPy/SciPy complex FP64, the MATLAB functions of the reference's hot path.  Every
function names the reference file:line it follows.  The reference is MATLAB, which is absent here
and on the GPU hosts, so it cannot be executed; each function is pinned instead by closed forms,
brute-force dense matrices on tiny grids, invariants and known-answer tests (``tests/test_oracle_*``).

Parity status per module:
  operator.py   pinned (closed-form dyadic, trace identity, dense brute force of the embedding)
  mvp.py        pinned (dense brute-force matrix on tiny grids, adjoint identity)
  circulant.py  pinned (textbook two-level Chan projection built independently, dense inverses)
  krylov.py     bicgstab/gmres: the MATLAB built-ins are closed source; the restatement follows
                their documented semantics and is pinned by exact-arithmetic KATs (identity,
                low-rank updates, few distinct eigenvalues) and by numpy.linalg.solve.  Iteration
                counts against MATLAB itself are parity unpinned.  pgmres follows pgmres.m exactly.
  farfield.py   pinned by the optical theorem and lossless energy balance (no reference code).
"""
