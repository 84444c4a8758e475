"""rkstiff_amd: rkstiff's stepping hot path on MI355X (gfx950) HIP kernels.

Drop-in solver classes (``ETD35, ETD34, ETD4, ETD5, IF34, IF45DP, IF4``) with
the reference's constructors, ``step`` and ``evolve`` on device tensors, the
reference's ``SolverConfig``/``ETDConfig``/``SolverType``, built-in nonlinear
terms (``nonlinear``), grids (``grids``), the Hankel transform
(``transforms``), the batched ensemble stepper (``batch.BatchSolver``) and its
learned step-size proposer (``learned``).  Compute runs in librkstiff_amd.so only.
"""
from .config import ETDConfig, SolverConfig, SolverType
from .nonlinear import (AdvectiveNL, AllenCahnNL, CubicNL, allen_cahn_ops, burgers_ops, kdv_ops, ks_ops,
                        nls_ops)
from .solvers import (ETD4, ETD5, ETD34, ETD35, IF4, IF34, IF45DP, BaseSolverAS, BaseSolverCS)
from .transforms import HankelTransform

__all__ = [
    "ETD35", "ETD34", "ETD4", "ETD5", "IF34", "IF45DP", "IF4", "BaseSolverAS", "BaseSolverCS",
    "SolverConfig", "ETDConfig", "SolverType", "AdvectiveNL", "CubicNL", "AllenCahnNL", "ks_ops",
    "kdv_ops", "burgers_ops", "nls_ops", "allen_cahn_ops", "HankelTransform",
]
