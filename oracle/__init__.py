"""CPU float64 oracle for the torchls block-sparse normal-equation path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline -- never as part of the product path
(``torchls_amd`` never imports it).

Every function restates the reference algorithm in numpy and cites the
reference file:line it follows.  How each function is pinned (golden vectors,
closed forms, invariants, brute force) is listed in ``PINS`` below; a function
whose entry says "parity unpinned" has no independent check.
"""

PINS = {
    "sparse.matvec/rmatvec/normal_matvec": "dense J from to_dense (closed form), adjointness, brute-force loops",
    "sparse.squared_matvec_norm": "equals ||dense J v||^2",
    "sparse.diagonal_blocks": "diagonal blocks of dense J^T J incl. aliased slots",
    "sparse.scale_columns": "equals dense J diag(s)",
    "precond.block_jacobi": "numpy.linalg.solve per block; golden fixture from the reference",
    "pcg.pcg_reference": "golden fixtures (iterates, counts) from the reference torch pcg; textbook PCG",
    "pcg.pcg_batched": "equals pcg_reference iterates while no check triggers; dense solve at convergence",
    "schur": "dense closed form S = A_cc - A_cp A_pp^-1 A_pc; Schur solve == full solve",
    "losses.huber": "closed form; golden fixture of the reference HuberLoss values/_reduction",
    "lm.solve": "golden LM trajectories (cost/lambda histories) from the reference solve()",
    "cholesky": "numpy.linalg.cholesky / LAPACK potrf on the same matrix",
}
