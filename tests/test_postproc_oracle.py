"""The steps either side of the Q map, in the oracle (oracle/curl.py,
oracle/qplane.py), pinned without the reference:

* the notebooks' index-space curl (demo/Q_RK4.ipynb cell [11]): central
  differences and the one-sided second-order z = 0 rule are exact for
  quadratic fields, so the curl of a quadratic B is known on every slice the
  notebook writes, and zero elsewhere;
* slog10 Q (demo/Q_RKF45CuPy.ipynb cells [13]-[14]): special values and the
  sign convention worked out by hand;
* the mixed method's selection (demo/Q_RKF45MixCuPyDiffScott.ipynb cell [9]).
"""
import numpy as np

from oracle import curl, qplane


def _quadratic_field(shape, seed):
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    X = [i, j, k]
    comps, grads = [], []
    for _ in range(3):
        a = rng.normal(size=3)
        Q = rng.normal(size=(3, 3)) * 0.1
        Q = Q + Q.T
        comps.append(sum(a[p] * X[p] for p in range(3)) +
                     0.5 * sum(Q[p, q] * X[p] * X[q] for p in range(3) for q in range(3)))
        grads.append([a[p] + sum(Q[p, q] * X[q] for q in range(3)) for p in range(3)])
    return comps, grads


def test_curl_index_exact_for_quadratics():
    shape = (7, 6, 5)
    (bx, by, bz), g = _quadratic_field(shape, 3)
    cx, cy, cz = curl.curl_index(bx, by, bz)
    want_x = g[2][1] - g[1][2]
    want_y = g[0][2] - g[2][0]
    want_z = g[1][0] - g[0][1]
    sx = (slice(None), slice(1, -1), slice(1, -1))
    sy = (slice(1, -1), slice(None), slice(1, -1))
    sz = (slice(1, -1), slice(1, -1), slice(None))
    np.testing.assert_allclose(cx[sx], want_x[sx], atol=1e-12)
    np.testing.assert_allclose(cy[sy], want_y[sy], atol=1e-12)
    np.testing.assert_allclose(cz[sz], want_z[sz], atol=1e-12)
    z0 = (slice(1, -1), slice(1, -1), 0)
    np.testing.assert_allclose(cx[z0], want_x[z0], atol=1e-12)
    np.testing.assert_allclose(cy[z0], want_y[z0], atol=1e-12)
    # untouched entries stay zero: curl_x on the y faces, the top z layer...
    assert np.all(cx[:, 0, :] == 0) and np.all(cx[:, -1, :] == 0) and np.all(cx[:, :, -1] == 0)
    assert np.all(cy[0, :, :] == 0) and np.all(cy[:, :, -1] == 0)
    assert np.all(cz[0, :, :] == 0) and np.all(cz[:, 0, :] == 0)
    # curl_x at z = 0 on the x faces is never written
    assert np.all(cx[0, :, 0] == 0) and np.all(cx[-1, :, 0] == 0)


def test_slog10q_by_hand():
    q = np.array([10, 100, 1, 0, np.nan, np.inf, -1, 100], np.float32)
    bn = np.array([2, -3, 0, 5, 1, -1, 1, np.nan], np.float32)
    got = qplane.slog10q(q, bn)
    want = np.array([1, -2, 0, -np.inf, np.nan, -np.inf, np.nan, np.nan], np.float32)
    # NumPy's float32 log10 is within an ulp (log10(100) gives 2 - 1 ulp here)
    np.testing.assert_allclose(got, want, rtol=2.5e-7, atol=0)
    flag = np.array([5, 6, 1, 42, 0, 7, 3, 2], np.int32)
    got = qplane.slog10q(q, bn, flag)
    want = np.array([1, -2, 0, np.nan, np.nan, np.nan, np.nan, np.nan], np.float32)  # inf * 0 = NaN
    np.testing.assert_allclose(got, want, rtol=2.5e-7, atol=0)
    assert got.dtype == np.float32


def test_bad_pixel_selection():
    b_flag = np.array([0, 1, 0, 1, 0, 0], np.int32)
    q = np.array([2.0, 3.0, np.nan, np.nan, 0.0, 1.0], np.float32)
    np.testing.assert_array_equal(qplane.bad_pixels(b_flag, q), [1, 2, 3])
