"""The steps either side of the Q map on the GPU (fastqsl_amd.postproc,
qmap.mixed_qmap) against the oracle's NumPy restatements:

* the notebooks' index-space curl: bit for bit, float32 and float64;
* slog10 Q: NaN / inf / zero placement and signs identical to NumPy, finite
  values within three float32 ulps of the float64 value;
* the mixed method's bad-pixel selection: the same pixels in the same
  order with their seeds; the fix-up writes the Scott tracer's q_perp for
  exactly those pixels and leaves every other pixel alone.
"""
import numpy as np
import pytest

from fastqsl_amd import synth
from oracle import curl, qplane

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
def test_curl_index_bit_exact(fq, dtype):
    from fastqsl_amd import postproc
    rng = np.random.default_rng(41)
    B = [rng.normal(size=(17, 13, 11)).astype(dtype) for _ in range(3)]
    got = postproc.curl_index(*B)
    want = curl.curl_index(*B)
    for g, w in zip(got, want):
        assert g.dtype == dtype and g.shape == w.shape
        np.testing.assert_array_equal(g, w)


def test_slog10q(fq):
    from fastqsl_amd import postproc
    rng = np.random.default_rng(42)
    n = 100000
    q = np.exp(rng.uniform(-2, 20, n)).astype(np.float32)
    q[:50] = [0, np.nan, np.inf, -1, 1] * 10
    bn = rng.normal(size=n).astype(np.float32)
    bn[50:80] = [0, np.nan, -0.0] * 10
    flag = rng.integers(-1, 8, n).astype(np.int32)
    flag[::97] = 42
    for f in (None, flag):
        got = postproc.slog10q(q, bn, f)
        want = qplane.slog10q(q, bn, f)
        np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
        np.testing.assert_array_equal(np.isinf(got), np.isinf(want))
        fin = np.isfinite(want)
        np.testing.assert_array_equal(np.sign(got[fin]), np.sign(want[fin]))
        # finite values: within 3 float32 ulps of log10 taken in float64 (the
        # device log10f and NumPy's are each within about an ulp of it)
        with np.errstate(divide="ignore", invalid="ignore"):
            exact = np.log10(q.astype(np.float64)) * np.sign(bn.astype(np.float64))
            if f is not None:
                exact = exact * (f <= 6)
        np.testing.assert_allclose(got[fin], exact[fin], rtol=3.6e-7, atol=1e-30)


def test_bad_pixels_and_fixup(fq):
    import torch
    from fastqsl_amd import postproc
    field = synth.quadrupole_64(dtype=np.float64, device="cpu", n=48)
    seeds, shape = synth.plane_seeds_z(60, 50, (1, 46), (1, 46), z=0.0)
    n = seeds.shape[0]
    rng = np.random.default_rng(43)
    q_all = rng.uniform(1, 100, n).astype(np.float32)
    q_all[rng.random(n) < 0.05] = np.nan
    b_flag = (rng.random(n) < 0.07).astype(np.int32)
    x, y, z = (np.ascontiguousarray(seeds[:, i]) for i in range(3))
    bx, by, bz, pix = postproc.bad_pixels(b_flag, q_all, x, y, z)
    want = qplane.bad_pixels(b_flag, q_all)
    np.testing.assert_array_equal(pix.cpu().numpy(), want)
    np.testing.assert_array_equal(bx.cpu().numpy(), x[want])
    np.testing.assert_array_equal(bz.cpu().numpy(), z[want])
    with fq.Field(*field) as f:
        qa = torch.as_tensor(q_all).cuda()
        m = postproc.scott_fixup(f, qa, b_flag, x, y, z, method="rkf45", step=0.125, tol=1e-6)
        assert m == want.size
        ref = fq.trace_scott(f, x[want], y[want], z[want], method="rkf45", step=0.125, tol=1e-6,
                             outputs=("q_perp",))
        torch.cuda.synchronize()
        got = qa.cpu().numpy()
        keep = np.ones(n, bool)
        keep[want] = False
        np.testing.assert_array_equal(got[keep], q_all[keep])
        np.testing.assert_array_equal(got[want], ref["q_perp"].cpu().numpy())


def test_mixed_qmap_replaces_flagged_pixels(fq):
    import torch
    from fastqsl_amd import qmap
    field = synth.quadrupole_64(dtype=np.float64, device="cpu", n=48)
    seeds, shape = synth.plane_seeds_z(80, 64, (2, 45), (2, 45), z=0.5)
    stride = 43.0 / 79
    with fq.Field(*field) as f:
        plain = qmap.neighbour_qmap(f, seeds, shape, stride, step=0.125, tol_coef=1.0)
        before = plain["q_all"].cpu().numpy().copy()
        bf = plain["traces"]["b_flag"].cpu().numpy()
        mixed = qmap.mixed_qmap(f, seeds, shape, stride, step=0.125, tol_coef=1.0)
        torch.cuda.synchronize()
        after = mixed["q_all"].cpu().numpy()
    bad = (bf + np.isnan(before)) > 0.5
    assert mixed["n_scott"] == int(bad.sum()) and bad.sum() > 0
    np.testing.assert_array_equal(after[~bad], before[~bad])
