"""Twist number and current density on the GPU.

* curl_of_B_grid / diff_3point / copy_boundary (FastQSL.py:120-163): bit for
  bit the NumPy restatement, float32 and float64 fields, non-uniform axes;
* J = alpha B: the Scott twist is alpha * length / 4 pi (the identity holds
  step by step and for the landing term);
* Scott and plain (RK4, RKF45) twist against the FP64 oracle with J from the
  grid curl, both precisions;
* the reference-shaped kernels with curB_flag set give the handle API's twist
  bit for bit;
* a uniform-twist Gold-Hoyle rope: 1.5 turns / sqrt(1 + T^2 r^2) near the axis.
"""
import math

import numpy as np
import pytest

from fastqsl_amd import synth
from oracle import curl
from tests import parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd
    fastqsl_amd.lib()
    return fastqsl_amd


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_curl_helpers_bit_exact(fq, dtype):
    from fastqsl_amd import FastQSL
    rng = np.random.default_rng(31)
    shape = (13, 11, 9)
    B = [rng.normal(size=shape).astype(dtype) for _ in range(3)]
    ax = [np.cumsum(rng.uniform(0.5, 1.5, n)) for n in shape]
    got = FastQSL.curl_of_B_grid(*B, *ax)
    want = curl.curl_of_B_grid(*B, *ax)
    for g, w in zip(got, want):
        assert g.dtype == np.float64 and g.shape == shape
        np.testing.assert_array_equal(g, w)
    for axis in range(3):
        np.testing.assert_array_equal(FastQSL.diff_3point(B[axis], ax[axis], axis),
                                      curl.diff_3point(B[axis], ax[axis], axis))
    a = rng.normal(size=shape)
    np.testing.assert_array_equal(FastQSL.copy_boundary(a.copy()), curl.copy_boundary(a.copy()))


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-12), ("fp32c", 2e-6)])
def test_scott_twist_of_constant_alpha(fq, precision, rtol):
    import torch
    f = synth.multipole((48, 40, 36), seed=4, dtype=np.float64, device="cpu")
    alpha = 0.3
    rng = np.random.default_rng(32)
    seeds = np.column_stack([rng.uniform(0, 47, 2000), rng.uniform(0, 39, 2000),
                             np.r_[np.zeros(1500), rng.uniform(0, 35, 500)]]).astype(np.float32)
    outs = ("q_0", "flag_start", "flag_end", "line_len", "twist")
    with fq.Field(*f, current=tuple(alpha * b for b in f)) as field:
        res = {}
        for sched in ("persistent", "per_line"):
            r = fq.trace_scott(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", step=0.25,
                               tol=1e-6, precision=precision, schedule=sched, outputs=outs)
            torch.cuda.synchronize()
            res[sched] = {k: r[k].cpu().numpy() for k in outs}
    r = res["persistent"]
    np.testing.assert_allclose(r["twist"], alpha * r["line_len"] / (4 * math.pi), rtol=rtol, atol=1e-12)
    for k in outs:
        assert np.array_equal(res["per_line"][k], r[k], equal_nan=True), k


def _curl_field(shape, seed):
    f = synth.multipole(shape, seed=seed, dtype=np.float64, device="cpu")
    return f, curl.curl_of_B_grid(*f, *(np.arange(n, dtype=np.float64) for n in shape))


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_scott_twist_matches_oracle(fq, oracle, precision):
    import torch
    f, cur = _curl_field((48, 40, 36), 6)
    rng = np.random.default_rng(33)
    seeds = np.column_stack([rng.uniform(2, 45, 1500), rng.uniform(2, 37, 1500),
                             np.r_[np.zeros(1000), rng.uniform(0, 30, 500)]]).astype(np.float32)
    ref = oracle.trace_scott(f, seeds, method="rk4", s0=0.25, current=cur)
    refp = [oracle.trace_scott(f, seeds.astype(np.float64) + sg * np.array([1e-6, -1e-6, 0.0]), method="rk4",
                               s0=0.25, current=cur) for sg in (1.0, -1.0)]
    cond = parity.conditioned(ref, *refp)
    for p in refp:
        cond &= np.abs(p["twist"] - ref["twist"]) <= 1e-6 * (1 + np.abs(ref["twist"]))
    with fq.Field(*f, current=cur) as field:
        r = fq.trace_scott(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rk4", step=0.25,
                           precision=precision, outputs=("flag_start", "flag_end", "twist", "q_0"))
        torch.cuda.synchronize()
        r = {k: v.cpu().numpy() for k, v in r.items() if k != "_seeds"}
    ok = cond & parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
    print(precision, "conditioned", cond.mean(), "ok", ok.mean())
    assert ok.mean() > 0.95
    rel = np.abs(r["twist"] - ref["twist"])[ok] / np.maximum(np.abs(ref["twist"]), 1e-3)[ok]
    assert np.quantile(rel, 0.999) < (1e-8 if precision == "fp64" else 5e-4)


@pytest.mark.parametrize("method", ["rk4", "rkf45"])
def test_plain_twist_matches_oracle(fq, oracle, method):
    import torch
    f, cur = _curl_field((48, 40, 36), 7)
    rng = np.random.default_rng(34)
    seeds = np.column_stack([rng.uniform(2, 45, 800), rng.uniform(2, 37, 800),
                             rng.uniform(0.5, 30, 800)]).astype(np.float32)
    kw = dict(method=method, s0=0.25, tol_coef=1e-3)
    ref = oracle.trace_plain(f, seeds, current=cur, **kw)
    refp = [oracle.trace_plain(f, seeds.astype(np.float64) + sg * np.array([1e-6, -1e-6, 0.0]), current=cur, **kw)
            for sg in (1.0, -1.0)]
    with fq.Field(*f, current=cur) as field:
        r = fq.trace_plain(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method=method, step=0.25,
                           tol_coef=1e-3, precision="fp64", outputs=("flag_start", "flag_end", "twist"))
        torch.cuda.synchronize()
        r = {k: v.cpu().numpy() for k, v in r.items() if k != "_seeds"}
    cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
    rel = lambda a: np.abs(a - ref["twist"]) / np.maximum(np.abs(ref["twist"]), 1e-3)
    if method == "rk4":   # deterministic: line by line on well-conditioned lines
        cond = parity.conditioned(ref, *refp, key="end")
        for p in refp:
            cond &= np.abs(p["twist"] - ref["twist"]) <= 1e-6 * (1 + np.abs(ref["twist"]))
        ok = cond & cls
        assert ok.mean() > 0.9
        assert np.quantile(rel(r["twist"])[ok], 0.99) < 1e-8
    else:
        # RKF45 step sequences fork on round-off even in FP64: the GPU's twist
        # is as close to the oracle's as the oracle is to itself under a
        # 1e-6-cell seed move
        self_spread = np.maximum(rel(refp[0]["twist"]), rel(refp[1]["twist"]))
        both = cls & parity.class_match(refp[0]["flag_start"], refp[0]["flag_end"], ref["flag_start"],
                                        ref["flag_end"])
        assert cls.mean() > 0.97
        for q in (0.5, 0.9):
            assert np.quantile(rel(r["twist"])[both], q) <= 1.5 * np.quantile(self_spread[both], q) + 1e-9


def test_dropin_twist_matches_handle_api(fq):
    import torch
    f, cur = _curl_field((40, 36, 30), 8)
    rng = np.random.default_rng(35)
    n = 1500
    seeds = np.column_stack([rng.uniform(1, 38, n), rng.uniform(1, 34, n), np.zeros(n)]).astype(np.float32)
    bx, by, bz = (np.asfortranarray(b.astype(np.float32)) for b in f)
    jx, jy, jz = (np.asfortranarray(c.astype(np.float32)) for c in cur)
    with fq.Field(*f, layouts=("soa32", "soa64"), current=(jx, jy, jz)) as field:
        ref = fq.trace_scott(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", step=0.25,
                             tol=1e-3 * float(np.float32(1e-3)), outputs=("q_0", "line_len", "twist"))
        refp = fq.trace_plain(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rk4", step=0.25,
                              outputs=("line_len", "twist"))
        torch.cuda.synchronize()
        ref = {k: v.cpu().numpy() for k, v in ref.items() if k != "_seeds"}
        refp = {k: v.cpu().numpy() for k, v in refp.items() if k != "_seeds"}
    f32 = lambda: np.zeros(n, np.float32)
    head = [bx, by, bz, np.array(bx.shape, np.int32), jx, jy, jz]
    tail = [np.ascontiguousarray(seeds[:, 0]), np.ascontiguousarray(seeds[:, 1]),
            np.ascontiguousarray(seeds[:, 2]), np.array([0, 0, 1], np.float32)]
    tw = np.zeros(n, np.float64)
    ll = np.zeros(n, np.float64)
    q0 = f32()
    args = head + [tw, np.array([True])] + tail + [f32(), f32(), f32(), np.zeros(n, np.int32), f32(), f32(), f32(),
                                                   np.zeros(n, np.int32), f32(), f32(), f32(), q0, f32(),
                                                   np.array([0.25], np.float32), np.array([n], np.uint64), ll,
                                                   np.array([1e-3], np.float32)]
    fq.kernels.compileTraceBlineScott()((16, 1), (256, 1, 1), tuple(args))
    np.testing.assert_array_equal(tw, ref["twist"])
    np.testing.assert_array_equal(ll, ref["line_len"])
    np.testing.assert_array_equal(q0, ref["q_0"])
    twp = np.zeros(n, np.float64)
    llp = np.zeros(n, np.float64)
    argsp = head + [twp, np.array([True])] + tail + [f32(), f32(), f32(), np.zeros(n, np.int32), f32(), f32(),
                                                     f32(), np.zeros(n, np.int32), f32(), f32(), f32(),
                                                     np.zeros(n, np.int32), f32(), f32(), f32(), f32(), f32(),
                                                     f32(), np.array([0.25], np.float32),
                                                     np.array([n], np.uint64), llp]
    fq.kernels.compileTraceBlineRK4()((16, 1), (256, 1, 1), tuple(argsp))
    np.testing.assert_array_equal(twp, refp["twist"])
    np.testing.assert_array_equal(llp, refp["line_len"])


def test_gold_hoyle_rope_twist(fq):
    import torch
    from fastqsl_amd import FastQSL
    shape = (129, 65, 65)
    turns, L = 1.5, 4.0
    f = synth.rope_arcade(shape=shape, dtype=np.float64, device="cpu", turns=turns, bipole_q=0.0)
    cur = FastQSL.curl_of_B_grid(*f, *(np.arange(n, dtype=np.float64) for n in shape))
    h = L / (shape[0] - 1)
    yc, zc = (shape[1] - 1) / 2, 0.4 / h
    rng = np.random.default_rng(36)
    rad = rng.uniform(0, 0.08, 200) / h
    ph = rng.uniform(0, 2 * math.pi, 200)
    seeds = np.column_stack([np.full(200, 64.0), yc + rad * np.cos(ph), zc + rad * np.sin(ph)]).astype(np.float32)
    T = 2 * math.pi * turns / L
    expect = turns / np.sqrt(1 + (T * rad * h) ** 2)
    with fq.Field(*f, current=cur) as field:
        for prec in ("fp64", "fp32c"):
            r = fq.trace_scott(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", step=0.25, tol=1e-7,
                               precision=prec, outputs=("flag_start", "flag_end", "twist"))
            torch.cuda.synchronize()
            assert np.all(r["flag_start"].cpu().numpy() == 1) and np.all(r["flag_end"].cpu().numpy() == 2)
            np.testing.assert_allclose(r["twist"].cpu().numpy(), expect, rtol=4e-3)
