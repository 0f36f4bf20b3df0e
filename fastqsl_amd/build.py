"""In-tree build of ``libfqsl.so`` for gfx950 with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

from ._lib import HERE, LIBPATH

ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "fqsl_kernels.hip")]
HEADERS = [os.path.join(HERE, "csrc", "fqsl_tracer.h"), os.path.join(HERE, "csrc", "fqsl_stretch.h"),
           os.path.join(ROOT, "include", "fqsl.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # keep packed FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    """Compile every HIP source into fastqsl_amd/libfqsl.so (always rebuilds)."""
    tmp = LIBPATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(HERE, "csrc"), *SOURCES, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l and l.strip())
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + err)
    if verbose and err:
        print(err, file=sys.stderr)
    os.replace(tmp, LIBPATH)
    return LIBPATH


if __name__ == "__main__":
    print(build(verbose=True))
