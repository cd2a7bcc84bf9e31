"""CPU oracle for the multigrid solve phase - TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker.  The product
(``amjax_amd``) never imports it.

``oracle.solve`` restates, in numpy/scipy, the algorithm of the reference's
``amjax/multilevel.py`` and ``amjax/relaxation/{relaxation,smoothing}.py``
line by line, plus the textbook preconditioned CG of
``jax.scipy.sparse.linalg.cg`` (jax >= 0.6.1, a third-party dependency of the
reference that is not installed here).
"""
