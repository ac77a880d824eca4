"""magudi_amd -- MI355X-native right-hand-side hot path of the magudi adjoint solver.

FP64 SBP/SAT compressible-flow RHS, its discrete adjoint, RK4 and HBM
checkpoint/recompute as hand-written HIP kernels for gfx950 behind the C ABI
in include/magudi_hip.h.  Python (this package) and Fortran (magudi_amd/fortran)
host code mirror magudi's Region / State / Patch / Solver interfaces.
"""
from . import cases  # noqa: F401

__all__ = ["cases", "capi", "solver"]


def __getattr__(name):
    if name in ("capi", "solver"):
        import importlib
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
