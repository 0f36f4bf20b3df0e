"""Writes the golden fixtures of this directory from the FP64 oracle.

    python tests/golden/make_golden.py

Each ``.npz`` holds the inputs (seeds; the field is regenerated from the
named synthetic generator and its arguments, stored as JSON in ``meta``) and
the oracle's outputs.  They pin the oracle against drift (tests/test_golden.py
regenerates them on the CPU) and give the GPU tests fixed expected values.
No reference-held vector with inputs exists for this path, so these are the
oracle's own -- themselves pinned by the closed-form tests of
tests/test_oracle.py and tests/test_qplane.py.  ``Q.fits`` is the reference
repository's own output file (no inputs), copied for the FITS format test.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from fastqsl_amd import synth  # noqa: E402
from oracle import oracle, qplane  # noqa: E402

CASES = {
    # name: (generator, generator kwargs, tracer, tracer kwargs, seed kind)
    "scott_quadrupole64_rk4": ("quadrupole_64", {}, "scott", dict(method="rk4", s0=0.25), "plane"),
    "scott_multipole64_rkf45": ("multipole", dict(shape=[64, 64, 64], seed=3), "scott",
                                dict(method="rkf45", s0=0.25, tol=1e-6), "plane"),
    "scott_multipole64_converged": ("multipole", dict(shape=[64, 64, 64], seed=3), "scott",
                                    dict(method="rkf45", s0=0.25, tol=1e-10), "plane"),
    "scott_rope_rk4_volume": ("rope_arcade", dict(shape=[48, 32, 32]), "scott",
                              dict(method="rk4", s0=0.25), "volume"),
    "plain_quadrupole64_rk4": ("quadrupole_64", {}, "plain", dict(method="rk4", s0=0.25), "lifted"),
}
N_LINES = 512


def field_of(gen, kw):
    kw = dict(kw)
    if "shape" in kw:
        kw["shape"] = tuple(kw["shape"])
    return getattr(synth, gen)(dtype=np.float64, device="cpu", **kw)


def seeds_of(kind, shape, rng):
    hi = np.array(shape, float) - 1
    s = rng.uniform(0, 1, size=(N_LINES, 3)) * hi
    if kind == "plane":
        s[:, 2] = 0.0
    elif kind == "lifted":
        s[:, 2] = 0.5
    return s.astype(np.float32)


def perturbed(trace, f, seeds, kw):
    return [trace(f, seeds.astype(np.float64) + sgn * np.array([1e-6, -1e-6, 0.0]), **kw)
            for sgn in (1.0, -1.0)]


def make(name):
    gen, gkw, tracer, tkw, kind = CASES[name]
    f = field_of(gen, gkw)
    # seeds depend on the field and the seed kind only, so the RKF45 and the
    # converged fixtures of one field trace the same lines
    rng = np.random.default_rng(sum(map(ord, gen + json.dumps(gkw, sort_keys=True) + kind)))
    seeds = seeds_of(kind, f[0].shape, rng)
    trace = oracle.trace_scott if tracer == "scott" else oracle.trace_plain
    r = trace(f, seeds, **tkw)
    pr = perturbed(trace, f, seeds, tkw)
    stable = np.ones(N_LINES, bool)
    for p in pr:
        stable &= (p["flag_start"] == r["flag_start"]) & (p["flag_end"] == r["flag_end"])
        if tracer == "scott":
            with np.errstate(invalid="ignore", divide="ignore"):
                d = np.abs(np.log10(p["q0"]) - np.log10(r["q0"]))
            stable &= (d < 1e-4) | (np.isnan(p["q0"]) & np.isnan(r["q0"])) | ((p["q0"] == 0) & (r["q0"] == 0))
        else:
            stable &= np.abs(p["end"] - r["end"]).max(axis=1) < 1e-3
            stable &= np.abs(p["start"] - r["start"]).max(axis=1) < 1e-3
    meta = dict(generator=gen, generator_kwargs=gkw, tracer=tracer, tracer_kwargs=tkw, seeds=kind)
    out = dict(meta=np.array(json.dumps(meta)), seeds=seeds, stable=stable)
    keep = ("start", "end", "flag_start", "flag_end", "length", "evals", "b_this")
    keep += ("q0", "qperp") if tracer == "scott" else ("b_start", "b_end", "b_flag")
    for k in keep:
        out[k] = r[k]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    return out


def make_qplane():
    """The neighbour-difference pipeline on the hyperbolic X-point field."""
    shape = (41, 41, 33)
    f = synth.hyperbolic(shape, k=0.03, b0=1.0)
    seeds, sh = synth.plane_seeds_z(12, 12, (17, 23), (17, 23), z=1e-4, dtype=np.float64)
    r = oracle.trace_plain(f, seeds, method="rk4", s0=0.125)
    g = lambda a, c=None: (a if c is None else a[:, c]).reshape(sh).astype(
        np.int32 if a.dtype == np.int32 else np.float32)
    args = (g(r["end"], 0), g(r["end"], 1), g(r["end"], 2), g(r["flag_end"]),
            g(r["start"], 0), g(r["start"], 1), g(r["start"], 2), g(r["flag_start"]),
            g(r["b_start"], 0), g(r["b_start"], 1), g(r["b_start"], 2),
            g(r["b_end"], 0), g(r["b_end"], 1), g(r["b_end"], 2),
            g(r["b_this"], 0), g(r["b_this"], 1), g(r["b_this"], 2), g(r["b_flag"]))
    q = qplane.qcalc_plane(*args, 6.0 / 11)
    names = ("x_end", "y_end", "z_end", "flag_end", "x_start", "y_start", "z_start", "flag_start",
             "bx_in", "by_in", "bz_in", "bx_out", "by_out", "bz_out", "bx_0", "by_0", "bz_0", "b_flag")
    np.savez_compressed(os.path.join(HERE, "qcalc_plane_hyperbolic.npz"), stride=np.float64(6.0 / 11),
                        q=q, **dict(zip(names, args)))


REFERENCE_QFITS = "/root/reference/Q.fits"


def copy_qfits():
    """Q.fits of the reference repository (q_0 and a BZ extension written by
    astropy in demo/Q_RKF45CuPyScott.ipynb cell [14]): a data file, kept here
    to check the FITS writer byte for byte against astropy's layout."""
    import shutil
    if os.path.exists(REFERENCE_QFITS):
        shutil.copyfile(REFERENCE_QFITS, os.path.join(HERE, "Q.fits"))
        return True
    return False


if __name__ == "__main__":
    print("Q.fits copied" if copy_qfits() else "Q.fits: reference file not present, kept as is")
    oracle.build()
    for name in CASES:
        o = make(name)
        print(name, "lines", N_LINES, "stable", int(o["stable"].sum()))
    make_qplane()
    print("qcalc_plane_hyperbolic")
