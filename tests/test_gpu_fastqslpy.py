"""The reference's own import name, driven the way its notebooks drive it.

``import fastqslpy.kernels`` / ``from fastqslpy import FastQSL`` (the
reference package, fastqslpy/__init__.py:1-3) with torch tensors where the
notebooks hold CuPy arrays:

* the neighbour-difference map of demo/Q_RKF45CuPy.ipynb cells [6]-[11]:
  TraceAllBline (RKF45), ResReshape, QCalcPlane, CookPseudoLine, the second
  TraceAllBline over the pseudo seeds, the notebook's own Qcut arithmetic
  and the merge -- against ``fastqsl_amd.qmap.neighbour_qmap`` on the same
  in-place float field, bit for bit;
* the Scott map of demo/Q_RKF45CuPyScott.ipynb cells [5]-[7]:
  TraceAllBlineScott with a 0-d device tolerance coefficient, ResReshape --
  against ``fastqsl_amd.trace_scott``, bit for bit.
"""
import numpy as np
import pytest

from fastqsl_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _field(torch, shape, seed):
    arrs = synth.multipole(shape, seed=seed, device="cpu")
    # x fastest in memory, as np.transpose(sav, (2, 1, 0)) leaves the notebooks' cubes
    dev = [torch.as_tensor(np.asfortranarray(b.astype(np.float32)).ravel(order="K")).cuda() for b in arrs]
    return arrs, dev


def test_neighbour_notebook_flow_through_fastqslpy(torch):
    import fastqslpy.kernels
    from fastqslpy import FastQSL
    import fastqsl_amd as fq
    shape = (72, 60, 48)
    arrs, (Bx_gpu, By_gpu, Bz_gpu) = _field(torch, shape, seed=5)
    TraceAllBline = fastqslpy.kernels.compileTraceBlineAdaptive()
    dummy = torch.zeros((1, 1), dtype=torch.float32, device="cuda")
    flag_twist = torch.tensor([False], device="cuda")
    BshapeN = torch.tensor(shape, dtype=torch.int32, device="cuda")
    interp_ratio = 4
    stride_step = 1 / interp_ratio
    x_range, y_range = [10, 50], [8, 40]
    x_i = torch.linspace(*x_range, int(interp_ratio * (x_range[1] - x_range[0])), dtype=torch.float32,
                         device="cuda")
    y_i = torch.linspace(*y_range, int(interp_ratio * (y_range[1] - y_range[0])), dtype=torch.float32,
                         device="cuda")
    y_arr, x_arr = torch.meshgrid(y_i, x_i, indexing="ij")      # cupy.meshgrid(x_i, y_i)
    xy_shape = tuple(x_arr.shape)
    x_inp = x_arr.flatten()
    y_inp = y_arr.flatten()
    z_inp = torch.zeros_like(x_inp) + 0.0001
    z32 = lambda: torch.zeros(x_inp.shape, dtype=torch.float32, device="cuda")
    (x_start, y_start, z_start, x_end, y_end, z_end, Bx_start, By_start, Bz_start, Bx_end, By_end, Bz_end,
     Bx_inp, By_inp, Bz_inp) = [z32() for _ in range(15)]
    B_flag, flag_start, flag_end = [torch.zeros(x_inp.shape, dtype=torch.int32, device="cuda") for _ in range(3)]
    line_len = torch.zeros(x_inp.shape, dtype=torch.float64, device="cuda")
    N = torch.tensor([x_inp.shape[0]], dtype=torch.int64, device="cuda")
    s_len = torch.tensor([1. / 8.], dtype=torch.float32, device="cuda")
    tol_coef = torch.tensor([1.], dtype=torch.float32, device="cuda")
    inp_norm = torch.tensor([0, 0, 1.], dtype=torch.float32, device="cuda")
    blck, grd = (128, 1, 1), (28, 1)
    TraceAllBline(blck, grd,
                  (Bx_gpu, By_gpu, Bz_gpu, BshapeN, dummy, dummy, dummy, dummy, flag_twist,
                   x_inp, y_inp, z_inp, inp_norm,
                   x_start, y_start, z_start, flag_start,
                   x_end, y_end, z_end, flag_end,
                   Bx_inp, By_inp, Bz_inp, B_flag,
                   Bx_start, By_start, Bz_start,
                   Bx_end, By_end, Bz_end,
                   s_len, N, line_len, tol_coef * 0.5))
    (x_end_arr, y_end_arr, z_end_arr, flag_end_arr,
     x_start_arr, y_start_arr, z_start_arr, flag_start_arr,
     Bx_in_arr, By_in_arr, Bz_in_arr,
     Bx_out_arr, By_out_arr, Bz_out_arr,
     Bx_0_arr, By_0_arr, Bz_0_arr, B_flag_arr) = FastQSL.ResReshape(
        xy_shape, x_end, y_end, z_end, flag_end, x_start, y_start, z_start, flag_start,
        Bx_start, By_start, Bz_start, Bx_end, By_end, Bz_end, Bx_inp, By_inp, Bz_inp, B_flag)
    Q = FastQSL.QCalcPlane(x_end_arr, y_end_arr, z_end_arr, flag_end_arr,
                           x_start_arr, y_start_arr, z_start_arr, flag_start_arr,
                           Bx_in_arr, By_in_arr, Bz_in_arr,
                           Bx_out_arr, By_out_arr, Bz_out_arr,
                           Bx_0_arr, By_0_arr, Bz_0_arr, B_flag_arr, stride_step)
    stride_this = np.float32(1. / interp_ratio)
    (cut_inp_x, cut_inp_y, cut_inp_z,
     cut_start_x, cut_start_y, cut_start_z, flag_cut_start,
     cut_end_x, cut_end_y, cut_end_z, flag_cut_end,
     Bx_inp_cut, By_inp_cut, Bz_inp_cut, B_flag_cut,
     Bx_start_cut, By_start_cut, Bz_start_cut,
     Bx_end_cut, By_end_cut, Bz_end_cut,
     N_cut, line_len_cut, Bz0_start, Bz0_end) = FastQSL.CookPseudoLine(
        x_end_arr, y_end_arr, z_end_arr, flag_end_arr,
        x_start_arr, y_start_arr, z_start_arr, flag_start_arr,
        Bx_in_arr, By_in_arr, Bz_in_arr,
        Bx_out_arr, By_out_arr, Bz_out_arr,
        Bz_0_arr, B_flag_arr, stride_this)
    assert int(N_cut[0]) > 0
    TraceAllBline(blck, grd,
                  (Bx_gpu, By_gpu, Bz_gpu, BshapeN, dummy, dummy, dummy, dummy, flag_twist,
                   cut_inp_x, cut_inp_y, cut_inp_z, inp_norm,
                   cut_start_x, cut_start_y, cut_start_z, flag_cut_start,
                   cut_end_x, cut_end_y, cut_end_z, flag_cut_end,
                   Bx_inp_cut, By_inp_cut, Bz_inp_cut, B_flag_cut,
                   Bx_start_cut, By_start_cut, Bz_start_cut,
                   Bx_end_cut, By_end_cut, Bz_end_cut,
                   s_len, N_cut, line_len_cut, tol_coef * .1))
    # the notebook's own Qcut arithmetic (cell [10]), here in torch
    X1, Y1, X2, Y2 = [torch.zeros(cut_inp_x.shape, dtype=torch.float32, device="cuda") for _ in range(4)]
    fdiv = lambda f: torch.div(f - 1, 2, rounding_mode="floor")
    iz1, iz2 = fdiv(flag_cut_start) == 2, fdiv(flag_cut_end) == 2
    iy1, iy2 = fdiv(flag_cut_start) == 1, fdiv(flag_cut_end) == 1
    ix1, ix2 = fdiv(flag_cut_start) == 0, fdiv(flag_cut_end) == 0
    X1[iz1] = cut_start_x[iz1]; Y1[iz1] = cut_start_y[iz1]
    X2[iz2] = cut_end_x[iz2]; Y2[iz2] = cut_end_y[iz2]
    X1[iy1] = cut_start_z[iy1]; Y1[iy1] = cut_start_x[iy1]
    X2[iy2] = cut_end_z[iy2]; Y2[iy2] = cut_end_x[iy2]
    X1[ix1] = cut_start_y[ix1]; Y1[ix1] = cut_start_z[ix1]
    X2[ix2] = cut_end_y[ix2]; Y2[ix2] = cut_end_z[ix2]
    dx2xc = X2[0::4] - X2[2::4]; dx2yc = X2[1::4] - X2[3::4]
    dy2xc = Y2[0::4] - Y2[2::4]; dy2yc = Y2[1::4] - Y2[3::4]
    dx1xc = X1[0::4] - X1[2::4]; dx1yc = X1[1::4] - X1[3::4]
    dy1xc = Y1[0::4] - Y1[2::4]; dy1yc = Y1[1::4] - Y1[3::4]
    a_cut = (dx2xc * dy1yc - dx2yc * dy1xc)
    b_cut = (dx2yc * dx1xc - dx2xc * dx1yc)
    c_cut = (dy2xc * dy1yc - dy2yc * dy1xc)
    d_cut = (dy2yc * dx1xc - dy2xc * dx1yc)
    bnr_cut = torch.abs(Bz0_end) / (torch.abs(Bz0_start)) * float(np.float32((1 / stride_this / 2) ** 4))
    Qcut = (a_cut ** 2 + b_cut ** 2 + c_cut ** 2 + d_cut ** 2) * bnr_cut
    Qcut[Qcut < 1.0] = 1.0
    Q_all = torch.zeros(x_end_arr.shape, dtype=torch.float32, device="cuda")
    Q_all[1:-1, 1:-1] = Q
    Q_all[B_flag_arr == 1] = Qcut
    Q_final = Q_all[1:-1, 1:-1]
    # the same map through the library's own pipeline
    with fq.Field(*arrs, layouts=("soa32", "soa64")) as f:
        seeds = torch.stack([x_inp, y_inp, z_inp], 1)
        # the drop-in reads tol_coef as float32: tol_coef * .1 is float32(0.1)
        mine = fq.qmap.neighbour_qmap(f, seeds, xy_shape, stride_step, method="rkf45", step=1. / 8.,
                                      tol_coef=0.5, pseudo_stride=float(stride_this),
                                      pseudo_tol_scale=float(np.float32(0.1)) / 0.5)
        torch.cuda.synchronize()
        assert int(mine["n_pseudo"]) == int(N_cut[0]) // 4
        np.testing.assert_array_equal(mine["q"].cpu().numpy(), Q_final.cpu().numpy())


def test_scott_notebook_flow_through_fastqslpy(torch):
    import fastqslpy
    import fastqsl_amd as fq
    FastQSL = fastqslpy.FastQSL
    shape = (72, 60, 48)
    arrs, (Bx_gpu, By_gpu, Bz_gpu) = _field(torch, shape, seed=6)
    TraceAllBlineScott = fastqslpy.kernels.compileTraceBlineScott()
    dummy = torch.zeros((1, 1), dtype=torch.float32, device="cuda")
    flag_twist = torch.tensor([False], device="cuda")
    BshapeN = torch.tensor(shape, dtype=torch.int32, device="cuda")
    interp_ratio = 2
    x_range, y_range = [12, 52], [10, 40]
    x_i = torch.linspace(*x_range, int(interp_ratio * (x_range[1] - x_range[0])), dtype=torch.float32,
                         device="cuda")
    y_i = torch.linspace(*y_range, int(interp_ratio * (y_range[1] - y_range[0])), dtype=torch.float32,
                         device="cuda")
    y_arr, x_arr = torch.meshgrid(y_i, x_i, indexing="ij")
    xy_shape = tuple(x_arr.shape)
    x_inp = x_arr.flatten()
    y_inp = y_arr.flatten()
    z_inp = torch.zeros_like(x_inp)
    (x_start, y_start, z_start, x_end, y_end, z_end, Bx_inp, By_inp, Bz_inp, q_0_gpu, q_perp_gpu) = [
        torch.zeros(x_inp.shape, dtype=torch.float32, device="cuda") for _ in range(11)]
    flag_start, flag_end = [torch.zeros(x_inp.shape, dtype=torch.int32, device="cuda") for _ in range(2)]
    line_len = torch.zeros(x_inp.shape, dtype=torch.float64, device="cuda")
    N = torch.tensor([x_inp.shape[0]], dtype=torch.int64, device="cuda")
    s_len = torch.tensor([1. / 8.], dtype=torch.float32, device="cuda")
    inp_norm = torch.tensor([0, 0, 1.], dtype=torch.float32, device="cuda")
    tol_coef_arr = torch.tensor([10 ** 2, 10 ** 1.5, 10 ** 1.0, 10 ** 0.5], dtype=torch.float32, device="cuda")
    real_tol = tol_coef_arr[3] / 1000                    # a 0-d device array, as in the notebook
    TraceAllBlineScott((128, 1, 1), (28, 1),
                       (Bx_gpu, By_gpu, Bz_gpu, BshapeN, dummy, dummy, dummy, dummy, flag_twist,
                        x_inp, y_inp, z_inp, inp_norm,
                        x_start, y_start, z_start, flag_start,
                        x_end, y_end, z_end, flag_end,
                        Bx_inp, By_inp, Bz_inp,
                        q_0_gpu, q_perp_gpu,
                        s_len, N, line_len, real_tol))
    (x_end_arr, y_end_arr, z_end_arr, flag_end_arr, x_start_arr, y_start_arr, z_start_arr, flag_start_arr,
     q_0_gpu_arr, q_perp_gpu_arr, Bx_0_arr, By_0_arr, Bz_0_arr) = FastQSL.ResReshape(
        xy_shape, x_end, y_end, z_end, flag_end, x_start, y_start, z_start, flag_start,
        q_0_gpu, q_perp_gpu, Bx_inp, By_inp, Bz_inp)
    torch.cuda.synchronize()
    assert q_perp_gpu_arr.shape == xy_shape
    with fq.Field(*arrs, layouts=("soa32", "soa64")) as f:
        tol = 1e-3 * float(real_tol.cpu())                # the drop-in's tolerance: 1e-3 x tol_coef
        r = fq.trace_scott(f, x_inp, y_inp, z_inp, method="rkf45", step=1. / 8., tol=tol)
        torch.cuda.synchronize()
        for mine, theirs in ((r["q_perp"], q_perp_gpu), (r["q_0"], q_0_gpu), (r["flag_end"], flag_end),
                             (r["end_x"], x_end), (r["start_z"], z_start), (r["b_this_z"], Bz_inp),
                             (r["line_len"], line_len)):
            np.testing.assert_array_equal(mine.cpu().numpy(), theirs.cpu().numpy())
    assert int(((flag_end >= 1) & (flag_end <= 6)).sum()) > 0.9 * x_inp.numel()
