"""CPU oracle for the AutoPDEx assemble-and-solve path (FP64 numpy/scipy).

TEST INFRASTRUCTURE ONLY.  This package is a plain restatement of the
reference algorithm (AutoPDEx 1.1.4) and is used purely as the checker:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product package ``autopdex_amd`` never
imports, calls or links anything from here, and no product code path falls
back to it.

Modules
-------
mesh      structured meshes, Dirichlet masks (psdf polygon, planes), Gauss rules
elements  Q1 / Hex8 shape functions and analytic element residuals/tangents
assembly  COO-with-duplicates assembly, canonical CSR, Dirichlet reduction
newton    Newton loop with the reference semantics, SuperLU and Jacobi-PCG
unet      numpy U-Net forward/backward used to check the learned preconditioner

Every function cites the reference file:line it restates.  How each function
is pinned (golden value, closed form, finite differences, independent
autograd) is listed in the docstring and exercised in ``tests/``.
"""
