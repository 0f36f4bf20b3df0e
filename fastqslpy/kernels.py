"""fastqslpy.kernels (fastqslpy/kernels.py:10-28 of peijin94/FastQSL): the
compile* functions return objects called as ``k(grid, block, args)`` with the
reference kernels' argument tuples; grid and block are accepted and ignored.
See ``fastqsl_amd.kernels`` for the argument conventions."""
from fastqsl_amd.kernels import (compileTraceBlineAdaptive, compileTraceBlineAdaptiveStretch,
                                 compileTraceBlineRK4, compileTraceBlineScott)

__all__ = ["compileTraceBlineAdaptive", "compileTraceBlineRK4", "compileTraceBlineScott",
           "compileTraceBlineAdaptiveStretch"]
