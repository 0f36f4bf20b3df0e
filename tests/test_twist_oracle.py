"""Twist number along field lines in the oracle (TraceBline.cu:325-365,
TraceBlineAdaptive.cu:373-384, TraceBlineScott.cu:423-457) and the
current-density restatement (FastQSL.py:120-163).

* J = alpha B makes J.B/|B|^2 = alpha at every point, so the Scott twist --
  step terms and the landing term alike -- is alpha * (line length) / 4 pi
  exactly; the plain RK4 tracer's landing term uses J.b with b normalised,
  so the same identity holds there in a field of unit magnitude.
* A uniform-twist Gold-Hoyle rope with J from the grid curl: field lines near
  the axis carry T L / (2 pi sqrt(1 + T^2 r^2)) turns (1.5 on the axis).
"""
import math

import numpy as np
import pytest

from fastqsl_amd import synth
from oracle import curl


@pytest.mark.parametrize("method,s0,tol", [("rk4", 0.25, 0.0), ("rkf45", 0.25, 1e-6)])
def test_scott_twist_of_constant_alpha(oracle, method, s0, tol):
    f = synth.multipole((40, 36, 32), seed=4, dtype=np.float64, device="cpu")
    alpha = 0.3
    cur = tuple(alpha * b for b in f)
    rng = np.random.default_rng(7)
    seeds = np.column_stack([rng.uniform(0, 39, 300), rng.uniform(0, 35, 300),
                             np.r_[np.zeros(200), rng.uniform(0, 31, 100)]])
    r = oracle.trace_scott(f, seeds, method=method, s0=s0, tol=tol, current=cur)
    assert set(np.unique(np.r_[r["flag_start"], r["flag_end"]])) >= {1, 3, 5}
    np.testing.assert_allclose(r["twist"], alpha * r["length"] / (4 * math.pi), rtol=1e-12, atol=1e-14)
    # without a current the twist is zero and nothing else changes but the
    # landing length rule (b at the last in-box state instead of raw B)
    r0 = oracle.trace_scott(f, seeds, method=method, s0=s0, tol=tol)
    assert np.all(r0["twist"] == 0)
    np.testing.assert_array_equal(r0["end"], r["end"])
    np.testing.assert_array_equal(r0["q0"], r["q0"])


def test_plain_rk4_twist_in_a_unit_field(oracle):
    b = np.array([0.3, 0.2, math.sqrt(1 - 0.13)])
    f = synth.uniform((24, 20, 30), b)
    alpha = -0.7
    cur = tuple(alpha * c for c in f)
    seeds = np.array([[5.0, 5.0, 3.0], [10.0, 7.0, 12.5], [2.0, 15.0, 20.0]])
    r = oracle.trace_plain(f, seeds, method="rk4", s0=0.5, current=cur)
    np.testing.assert_allclose(r["twist"], alpha * r["length"] / (4 * math.pi), rtol=1e-12)
    # RKF45: no landing term, so the twist covers the stepped length only
    r = oracle.trace_plain(f, seeds, method="rkf45", s0=0.5, current=cur)
    stepped = r["twist"] * 4 * math.pi / alpha
    assert np.all(stepped <= r["length"] + 1e-12) and np.all(stepped > 0.5 * r["length"])


def test_curl_restatement_on_a_linear_field():
    """The three-point difference is exact for linear data on any spacing, and
    the boundary copies the nearest interior value."""
    rng = np.random.default_rng(3)
    ax = [np.cumsum(rng.uniform(0.5, 1.5, n)) for n in (6, 7, 8)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    A = rng.normal(size=(3, 3))
    B = [A[i, 0] * X + A[i, 1] * Y + A[i, 2] * Z for i in range(3)]
    cx, cy, cz = curl.curl_of_B_grid(*B, *ax)
    want = (A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1])
    for c, w in zip((cx, cy, cz), want):
        np.testing.assert_allclose(c, w, atol=1e-12)
    # quadratic data: the non-uniform three-point rule is second order,
    # d/dx of x^2 at x_i is exactly x_{i-1} + x_{i+1} - ... not needed: check
    # against its definition at one node
    q = X ** 2
    d = curl.diff_3point(q, ax[0], 0)
    i = 2
    h0, h1 = ax[0][i] - ax[0][i - 1], ax[0][i + 1] - ax[0][i]
    expect = ((ax[0][i + 1] ** 2 - ax[0][i] ** 2) / h1 * h0 / (h0 + h1)
              + (ax[0][i] ** 2 - ax[0][i - 1] ** 2) / h0 * h1 / (h0 + h1))
    np.testing.assert_allclose(d[i - 1, 0, 0], expect, rtol=1e-14)


def test_gold_hoyle_twist(oracle):
    shape = (129, 65, 65)
    turns, L = 1.5, 4.0
    f = synth.rope_arcade(shape=shape, dtype=np.float64, device="cpu", turns=turns, bipole_q=0.0)
    cur = curl.curl_of_B_grid(*f, *(np.arange(n, dtype=np.float64) for n in shape))
    h = L / (shape[0] - 1)                       # cell size (cubic cells)
    yc, zc = (shape[1] - 1) / 2, 0.4 / h
    rng = np.random.default_rng(9)
    rad = rng.uniform(0, 0.08, 40) / h
    ph = rng.uniform(0, 2 * math.pi, 40)
    seeds = np.column_stack([np.full(40, 64.0), yc + rad * np.cos(ph), zc + rad * np.sin(ph)])
    r = oracle.trace_scott(f, seeds, method="rkf45", s0=0.25, tol=1e-7, current=cur)
    assert np.all(r["flag_start"] == 1) and np.all(r["flag_end"] == 2)
    T = 2 * math.pi * turns / L
    expect = turns / np.sqrt(1 + (T * rad * h) ** 2)
    np.testing.assert_allclose(r["twist"], expect, rtol=4e-3)
