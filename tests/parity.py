"""Parity metrics shared by the GPU tests (GPU result vs FP64 oracle)."""
import numpy as np

# Seeds the reference's drivers meet at the edges of a 32^3 box: below the
# surface, on faces and edges, on the strict y = 0 plane, outside the box,
# NaN, at the centre (a null of the b0 = 0 hyperbolic field), just above z = 0.
EDGE_SEEDS = np.array([[10.0, 10.0, -0.5], [0.0, 5.0, 0.0], [31.0, 12.0, 3.0], [5.0, 0.0, 2.0],
                       [15.5, 15.5, 4.0], [np.nan, 3.0, 3.0], [40.0, 3.0, 3.0], [31.0, 31.0, 31.0],
                       [0.0, 0.0, 0.0], [12.0, 12.0, 31.0], [3.0, 4.0, 1e-7]], np.float32)


def log_q_diff(q_gpu, q_ref):
    """|log10 q_gpu - log10 q_ref| per line; 0 where both are NaN or both 0,
    inf where only one is finite/positive."""
    a = np.asarray(q_gpu, np.float64)
    b = np.asarray(q_ref, np.float64)
    d = np.full(a.shape, np.inf)
    ok = np.isfinite(a) & np.isfinite(b) & (a > 0) & (b > 0)
    d[ok] = np.abs(np.log10(a[ok]) - np.log10(b[ok]))
    d[np.isnan(a) & np.isnan(b)] = 0.0
    d[(a == 0) & (b == 0)] = 0.0
    return d


def class_match(fs_a, fe_a, fs_b, fe_b):
    return (np.asarray(fs_a) == np.asarray(fs_b)) & (np.asarray(fe_a) == np.asarray(fe_b))


def conditioned(ref, *perturbed, thresh=1e-4, key="q0"):
    """Lines whose oracle result is stable under perturbations far below
    float32 resolution (seed moved by +-1e-6 cells, or tolerance scaled by
    1 + 1e-12): same classes and Q (or, with key="end", footpoints within
    1e-3 cells) in every perturbed run.  Lines on a separatrix through a null
    (Q -> infinity) or whose adaptive step sequence forks on round-off fail
    this."""
    ok = np.ones(np.asarray(ref["flag_start"]).shape, bool)
    for p in perturbed:
        ok &= class_match(ref["flag_start"], ref["flag_end"], p["flag_start"], p["flag_end"])
        if key == "q0":
            ok &= log_q_diff(ref["q0"], p["q0"]) < thresh
        else:
            ok &= np.abs(np.asarray(ref["end"]) - p["end"]).max(axis=1) < 1e-3
            ok &= np.abs(np.asarray(ref["start"]) - p["start"]).max(axis=1) < 1e-3
    return ok


def summary(d, cls, mask=None):
    if mask is not None:
        d = d[mask]
        cls = cls[mask]
    fin = d[np.isfinite(d)]
    return dict(n=int(d.size), class_match=float(cls.mean()) if cls.size else 1.0,
                frac_le_1e3=float(np.mean(d <= 1e-3)) if d.size else 1.0,
                p999=float(np.quantile(fin, 0.999)) if fin.size else 0.0,
                median=float(np.median(fin)) if fin.size else 0.0,
                max=float(fin.max()) if fin.size else 0.0)


def ulp_moved(seeds, hi, rng):
    """Seeds with x and y each moved by one float32 ulp, up or down at random
    (default_rng stream `rng`), kept inside [0, hi]; z unchanged.  The oracle
    traced from them measures how far round-off of float32 size alone moves
    each line's class and Q: the floor an FP32 path is held to."""
    s = np.array(seeds, np.float32, copy=True)
    for a in (0, 1):
        up = rng.random(s.shape[0]) < 0.5
        moved = np.where(up, np.nextafter(s[:, a], np.float32(np.inf)),
                         np.nextafter(s[:, a], np.float32(-np.inf)))
        s[:, a] = np.clip(moved, np.float32(0), np.float32(hi[a]))
    return s
