"""fastqsl_amd -- MI355X-native squashing-factor (Q) engine.

Field-line tracing with the Scott Jacobian and the neighbour-difference Q
method on gfx950, behind a C ABI (include/fqsl.h, libfqsl.so) with a Python
host side that keeps the ``fastqslpy`` entry points:

* ``fastqsl_amd.kernels``  -- compileTraceBlineAdaptive / compileTraceBlineRK4 /
  compileTraceBlineScott, callable as ``k(grid, block, args)``
* ``fastqsl_amd.FastQSL``  -- QCalcPlane, CookPseudoLine, ResReshape
* ``Field``, ``trace_scott``, ``trace_plain`` -- the direct API
"""
from ._lib import FqslError, lib, last_kernel_ms
from .field import Field
from .trace import trace_plain, trace_scott
from . import synth

__version__ = "0.1.0"

__all__ = ["Field", "trace_scott", "trace_plain", "FqslError", "lib", "last_kernel_ms", "synth",
           "kernels", "FastQSL"]


def __getattr__(name):
    # the fastqslpy mirror modules import torch lazily
    if name in ("kernels", "FastQSL", "qmap"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
