"""fastqslpy -- the reference package's import name on the MI355X library.

``import fastqslpy.kernels`` and ``import fastqslpy.FastQSL`` work unchanged
(fastqslpy/__init__.py:1-3 of peijin94/FastQSL): both modules are the
``fastqsl_amd`` host-side mirrors, which launch hand-written gfx950 kernels
through ``libfqsl.so`` and take NumPy arrays or ROCm torch tensors where the
reference took CuPy arrays.  ``MagTool`` (a NumPy toy line tracer outside the
Q engine) is not provided.
"""
from . import kernels
from . import FastQSL

__all__ = ["kernels", "FastQSL"]


def __getattr__(name):
    if name == "MagTool":
        raise AttributeError("fastqslpy.MagTool (the NumPy toy tracer) is not part of this library; "
                             "use fastqslpy.kernels or fastqsl_amd.trace_plain")
    raise AttributeError(name)
