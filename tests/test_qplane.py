"""Known-answer tests of the neighbour-difference Q assembly restatement
(oracle/qplane.py: QCalcPlane, FastQSL.py:65-117; CookPseudoLine,
FastQSL.py:4-57; Qcut and the merge, demo/Q_RKF45CuPy.ipynb cells [10]-[11]).

For an affine footpoint map X2 = M X1 between two faces, with Bn at the two
ends in the ratio fixed by flux conservation (|Bn_end / Bn_start| =
1 / |det M|), the squashing factor is the textbook ||M||_F^2 / |det M|.  The
tests feed such maps through each exit face's coordinate convention, so a
swapped difference, a wrong face axis, a wrong stride power or an inverted Bn
ratio fails them.
"""
import numpy as np
import pytest

from oracle import qplane

f32 = np.float32
# exit class -> (axis of the face normal, in-face (X, Y) axes) of the reference
FACE_AXES = {1: (0, 1, 2), 2: (0, 1, 2), 3: (1, 2, 0), 4: (1, 2, 0), 5: (2, 0, 1), 6: (2, 0, 1)}
# start class -> the two axes CookPseudoLine offsets, seeds 0/2 and 1/3
PSEUDO_AXES = {1: (1, 2), 2: (1, 2), 3: (0, 2), 4: (0, 2), 5: (0, 1), 6: (0, 1)}


def _place(X, Y, cls, level):
    """3-D point on the face of class `cls` with in-face coordinates (X, Y)."""
    n, a, b = FACE_AXES[cls]
    p = [None, None, None]
    p[a], p[b], p[n] = X, Y, np.full_like(X, level)
    return p


def _plane(M, stride, start_cls=5, end_cls=6, ny=7, nx=9, bn0=1.5):
    """Traces of a seed grid on the start face whose end footpoints are
    M @ (X1, Y1) + t on the end face; Bn at the start = bn0, at the end
    bn0 / |det M|, B at the seed plane = (0, 0, bn0)."""
    M = np.asarray(M, np.float64)
    xs = (np.arange(nx) * stride + 3.0)
    ys = (np.arange(ny) * stride + 2.0)
    X1, Y1 = np.meshgrid(xs, ys)                    # (ny, nx): rows = second in-face axis
    X2 = M[0, 0] * X1 + M[0, 1] * Y1 + 0.5
    Y2 = M[1, 0] * X1 + M[1, 1] * Y1 - 0.25
    s = [a.astype(f32) for a in _place(X1, Y1, start_cls, 0.0)]
    e = [a.astype(f32) for a in _place(X2, Y2, end_cls, 40.0)]
    fs = np.full((ny, nx), start_cls, np.int32)
    fe = np.full((ny, nx), end_cls, np.int32)
    bn_end = bn0 / abs(np.linalg.det(M))
    ones = np.ones((ny, nx), f32)
    return dict(x_end_arr=e[0], y_end_arr=e[1], z_end_arr=e[2], flag_end_arr=fe,
                x_start_arr=s[0], y_start_arr=s[1], z_start_arr=s[2], flag_start_arr=fs,
                Bx_in_arr=0 * ones, By_in_arr=0 * ones, Bz_in_arr=f32(bn0) * ones,
                Bx_out_arr=0 * ones, By_out_arr=0 * ones, Bz_out_arr=f32(bn_end) * ones,
                Bx_0_arr=0 * ones, By_0_arr=0 * ones, Bz_0_arr=f32(bn0) * ones,
                B_flag_arr=np.zeros((ny, nx), np.int32), stride_step=stride)


def _q_textbook(M):
    M = np.asarray(M, np.float64)
    return np.sum(M * M) / abs(np.linalg.det(M))


def _call(d):
    return qplane.qcalc_plane(d["x_end_arr"], d["y_end_arr"], d["z_end_arr"], d["flag_end_arr"],
                              d["x_start_arr"], d["y_start_arr"], d["z_start_arr"],
                              d["flag_start_arr"], d["Bx_in_arr"], d["By_in_arr"], d["Bz_in_arr"],
                              d["Bx_out_arr"], d["By_out_arr"], d["Bz_out_arr"],
                              d["Bx_0_arr"], d["By_0_arr"], d["Bz_0_arr"], d["B_flag_arr"],
                              d["stride_step"])


MAPS = [np.diag([3.0, 1 / 3.0]),                                   # Q = 9 + 1/9
        [[1.0, 0.0], [0.0, 1.0]],                                  # identity: Q = 2
        [[0.8, -1.3], [0.45, 0.9]],                                # general
        [[0.0, 2.0], [-0.7, 0.1]]]                                 # axes swapped


@pytest.mark.parametrize("M", MAPS)
@pytest.mark.parametrize("end_cls", [6, 2, 4, 3])
def test_qcalc_plane_affine_map(M, end_cls):
    d = _plane(M, 0.5, end_cls=end_cls)
    Q = _call(d)
    assert Q.shape == (5, 7) and Q.dtype == np.float32
    np.testing.assert_allclose(Q, _q_textbook(M), rtol=2e-5)


def test_qcalc_plane_rotation_invariance_and_stride():
    """Q is invariant under a rotation of the end face and independent of the
    pixel spacing: (1 / (2 stride))^4 cancels the stride of the differences."""
    M = np.array([[1.7, 0.2], [-0.4, 0.5]])
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    q = _q_textbook(M)
    for stride in (0.25, 1.0, 3.0):
        np.testing.assert_allclose(_call(_plane(M, stride)), q, rtol=2e-5)
        np.testing.assert_allclose(_call(_plane(R @ M, stride)), q, rtol=2e-5)


def test_qcalc_plane_nan_mask_clamp_and_oblique_plane():
    d = _plane(np.diag([0.1, 0.1]), 1.0)
    # ||M||^2 / |det| = 2 here, but an (unphysical) Bn ratio of 1 makes the raw
    # value 0.02: the reference clamps Q < 1 to 1
    d["Bz_out_arr"] = d["Bz_in_arr"].copy()
    np.testing.assert_array_equal(_call(d), 1.0)
    # one pixel with another end class poisons itself and its 4 neighbours
    d = _plane(np.eye(2), 1.0)
    d["flag_end_arr"][3, 4] = 4
    Q = _call(d)
    nan = np.isnan(Q)
    want = np.zeros_like(nan)
    for (r, c) in [(3, 4), (2, 4), (4, 4), (3, 3), (3, 5)]:
        want[r - 1, c - 1] = True
    np.testing.assert_array_equal(nan, want)
    np.testing.assert_allclose(Q[~nan], 2.0, rtol=1e-6)
    # Bn at the seed plane is B_0 . CrossBC_dir: an oblique B_0 with the same
    # normal component gives the same Q
    d = _plane(np.diag([2.0, 0.5]), 1.0)
    d["Bx_0_arr"] = np.full_like(d["Bx_0_arr"], 7.0)
    np.testing.assert_allclose(_call(d), 4.25, rtol=2e-5)


def _pseudo_case(M, cls_start, stride, bn0=2.0):
    """m pixels on the start face of class cls_start, traced by the affine map M."""
    rng = np.random.default_rng(3)
    m = 6
    X1, Y1 = rng.uniform(5, 20, m), rng.uniform(5, 20, m)
    level = 0.0 if cls_start % 2 else 30.0
    s = _place(X1, Y1, cls_start, level)
    bf = np.zeros((4, 5), np.int32)
    pix = np.sort(rng.choice(bf.size, m, replace=False))
    bf.reshape(-1)[pix] = 1
    full = lambda v: np.zeros(bf.size, f32)
    sx, sy, sz = full(0), full(0), full(0)
    sx[pix], sy[pix], sz[pix] = s[0], s[1], s[2]
    fl = np.zeros(bf.size, np.int32)
    fl[pix] = cls_start
    bzs, bze = full(0), full(0)
    bzs[pix] = bn0
    bze[pix] = bn0 / abs(np.linalg.det(M))
    sh = bf.shape
    return (sx.reshape(sh), sy.reshape(sh), sz.reshape(sh), fl.reshape(sh), bzs.reshape(sh),
            bze.reshape(sh), bf, pix, (X1, Y1))


@pytest.mark.parametrize("cls_start", [5, 3, 1, 6])
def test_pseudo_seeds_and_qcut(cls_start):
    M = np.array([[0.9, -0.6], [0.35, 1.4]])
    stride = 0.125
    sx, sy, sz, fl, bzs, bze, bf, pix, _ = _pseudo_case(M, cls_start, stride)
    seeds, b0s, b0e = qplane.pseudo_seeds(sx, sy, sz, fl, bzs, bze, bf, stride)
    m = len(pix)
    assert seeds.shape == (4 * m, 3) and seeds.dtype == np.float32
    n, a, b = FACE_AXES[cls_start]
    oa, ob = PSEUDO_AXES[cls_start]
    # order (+oa, +ob, -oa, -ob) around the start footpoint, the normal coordinate kept
    off = np.array([[1, 0], [0, 1], [-1, 0], [0, -1]], np.float64) * stride
    for i in range(m):
        for k in range(4):
            p0 = np.array([sx.reshape(-1)[pix[i]], sy.reshape(-1)[pix[i]], sz.reshape(-1)[pix[i]]])
            np.testing.assert_allclose(seeds[4 * i + k, oa], p0[oa] + off[k, 0], atol=1e-5)
            np.testing.assert_allclose(seeds[4 * i + k, ob], p0[ob] + off[k, 1], atol=1e-5)
            assert seeds[4 * i + k, n] == (0.0 if cls_start % 2 else 30.0)
    np.testing.assert_array_equal(b0s, bzs.reshape(-1)[pix])
    np.testing.assert_array_equal(b0e, bze.reshape(-1)[pix])
    # trace the pseudo seeds by the affine map onto the z max face
    Xs, Ys = seeds[:, a].astype(np.float64), seeds[:, b].astype(np.float64)
    e = _place(M[0, 0] * Xs + M[0, 1] * Ys, M[1, 0] * Xs + M[1, 1] * Ys, 6, 50.0)
    end = np.stack(e, axis=1).astype(f32)
    q = qplane.qcut(seeds, np.full(4 * m, cls_start, np.int32), end, np.full(4 * m, 6, np.int32),
                    b0s, b0e, stride)
    np.testing.assert_allclose(q, _q_textbook(M), rtol=2e-4)
    # merge: Qcut lands on the flagged pixels, the rest comes from QCalcPlane
    inner = np.full((bf.shape[0] - 2, bf.shape[1] - 2), 5.0, f32)
    Qf = qplane.merge(inner, q, bf)
    want = np.full(bf.shape, 0.0, f32)
    want[1:-1, 1:-1] = 5.0
    want.reshape(-1)[pix] = q
    np.testing.assert_array_equal(Qf, want[1:-1, 1:-1])


def test_footpoint_coordinates_of_other_classes_are_zero():
    x = np.array([1.0, 2.0, 3.0], f32)
    X, Y = qplane._footpoint_xy(x, x + 1, x + 2, np.array([0, 42, -1], np.int32))
    np.testing.assert_array_equal(X, 0.0)
    np.testing.assert_array_equal(Y, 0.0)
