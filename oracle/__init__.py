"""FP64 CPU oracle for the periodic inverse-homogenization inner loop.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``maple_amd``) imports,
links or calls this package.  It is used by ``tests/`` as the checker, by
``__graft_entry__.smoke()`` to check one small GPU invocation, and by the
``cpu_baseline`` leg of ``bench.py`` (timed on the host cores).

It restates, in numpy/scipy FP64, the algorithm of the reference
(xing-yuu/MAPLE, ``src/maple/core``):

* ``fe``        element matrices            (_utils.py:39-196)
* ``grid``      periodic hex8 grid, RHS      (main_IHD_v3.py:162-215, _utils.py:453-507)
* ``mg``        converged solve: Galerkin multigrid-preconditioned CG
                (EBE_GMG.py:202-430, 470-647; converged PCG with a
                weighted-Jacobi V-cycle and a dense coarsest solve)
* ``homog``     energy-form C_H / kappa_H and closed-form sensitivities
                (model.py:33-95; main_IHD_v3.py:688 at convergence)
* ``objective`` stiffness / shear / thermal terms (objective.py:11-251)
                plus the bulk-modulus extension
* ``oc``        periodic box filter and OC bisection (main_IHD_v3.py:342-380)
* ``driver``    the plain OC loop (main_IHD_v3.py:620-716)

Pinning status of every function is listed in each module header.
"""
