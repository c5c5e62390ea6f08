"""maple_amd: the periodic inverse-homogenization inner loop of MAPLE on MI355X.

Hand-written gfx950 HIP kernels behind a C ABI (include/maple_hip.h,
maple_amd/lib/libmaple_hip.so), driven from Python over PyTorch-ROCm device
memory and streams.

* ``CellBatch``            one context: a batch of periodic cells on one GPU
* ``IHD``                  drop-in for the reference's homogenization module
* ``OCDriver``             the plain OC loop on the device
* ``ConcurrentOC``         several OCDriver batches advancing side by side (own stream
                           and host thread each)
* ``optimize_inverse_homogenization``: drop-in for the reference's OC driver
  (symmetry projection, early stop, iteration log, summary, final density)
* ``avg_filter3d_x_periodic``, ``oc_update``, ``build_unitcell_batch``,
  ``assemble_F_fast``, ``assemble_F_fast_scalar``: reference-named drop-ins
"""
from .cells import CellBatch, hex8_matrices, max_levels
from .model import IHD, make_cfg
from .objective import (objective_from_spec, objective_requires, parse_objective_spec, shear_objective,
                        stiffness_objective, thermal_objective, bulk_objective)
from .driver import optimize_inverse_homogenization, parse_symmetry, symmetrize_vec
from .optim import (OCDriver, assemble_F_fast, assemble_F_fast_scalar, avg_filter3d_x_periodic,
                    build_unitcell_batch, oc_update, x0_sphere_void, x0_uniform_noise,
                    batched_workload, simplex_grid)
from .concurrent import ConcurrentOC

__all__ = [
    "CellBatch", "hex8_matrices", "max_levels", "IHD", "make_cfg", "objective_from_spec", "objective_requires",
    "parse_objective_spec", "shear_objective", "stiffness_objective", "thermal_objective", "bulk_objective",
    "OCDriver", "assemble_F_fast", "assemble_F_fast_scalar", "avg_filter3d_x_periodic", "build_unitcell_batch",
    "oc_update", "x0_sphere_void", "x0_uniform_noise", "batched_workload", "simplex_grid",
    "optimize_inverse_homogenization", "parse_symmetry", "symmetrize_vec", "ConcurrentOC",
]
