"""MI355X-native Kirchhoff-Love shell equilibrium path (gsStructuralAnalysis hot path).

The compute path is libksh.so (HIP kernels for gfx950 behind the C-ABI in include/ksh.h);
this package adds the geometry builders and a Python front-end.  Importing
:mod:`gsstructuralanalysis_amd.problem` requires the built library; there is no CPU
execution path.
"""
from . import geometry  # noqa: F401

__all__ = ["geometry", "problem"]
