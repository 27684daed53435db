"""MI355X-native PS-IPM hot path.

The preconditioned-CG solve of the proximal normal equations
``(A W A' + delta I) dy = r`` with an active-column sparse Cholesky
preconditioner, run by hand-written HIP kernels for gfx950 behind a C-ABI
(``libpsipm_hip.so``) and exposed through the same Python object interface as
the reference's ``SpectralStableColumnPreconditioner``.
"""
from . import lp, generators, ipm  # noqa: F401

__all__ = ["lp", "generators", "ipm"]
