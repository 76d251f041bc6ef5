"""staccato_amd -- MI355X (gfx950) frequency sweep for Krylov reduced-order models.

For every excitation frequency the library forms
A(w) = (1 + i eta) K + i w D - w^2 M and solves A(w) X = F by pivoted LU on the
GPU (plain complex128, or complex64 factors with FP64 refinement).  The C ABI
is include/staccato_sweep.h; this package is its Python host side.
"""
from .sweep import Handle, Plan, StcError, MODES  # noqa: F401
from . import synth  # noqa: F401
from . import coupling  # noqa: F401
from . import autotune  # noqa: F401

__all__ = ["Handle", "Plan", "StcError", "MODES", "synth", "coupling", "autotune"]
