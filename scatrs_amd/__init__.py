"""scatrs_amd: the size-distribution ensemble path of scatrs (homogeneous and layered spheres)
on AMD Instinct MI355X (gfx950), as hand-written HIP kernels behind a C ABI.

The Python layer mirrors the reference's package API for this path (scatrs.mie.efficiencies,
scatrs.Ensemble.compute and Mixture.compute) and, in _wire, the ensemble functions of its
compiled module (ensemble_parameters, ensemble_compute, ensemble_compute_batch). Ensembles are
integrated by the adaptive Gauss-Kronrod engine (adaptive) or on a fixed grid; device memory
and streams come from PyTorch-ROCm.
"""

from .errors import (InvalidInputError, NativeLibraryError, NotConvergedError,
                     OutsideValidatedDomainError, ScatrsError)

__all__ = [
    "InvalidInputError",
    "NativeLibraryError",
    "NotConvergedError",
    "OutsideValidatedDomainError",
    "ScatrsError",
]
