"""Field ingest and result files (fastqsl_amd.io).

* FITS: the reference repository's Q.fits (written by astropy in
  demo/Q_RKF45CuPyScott.ipynb cell [14]; tests/golden/Q.fits) reads back as
  q_0 (100 x 240) with a BZ extension, and writing those arrays again gives
  the same bytes; other dtypes and ranks round-trip.
* npz: the staggered layout of demo/npzinputRKF45CuPyTwist.ipynb cell [3].
* IDL save: scipy's reader is stood in for by the structure it returns (a
  dict of lower-case variables, or a structure array with upper-case fields
  and lower-case titles); the (2, 1, 0) transpose and the float32 cast are
  what is checked.  No .sav file with a field exists in the reference
  (B0.sav is not in the repository), so reading a real file is parity
  unpinned.
"""
import os

import numpy as np
import pytest

from fastqsl_amd import io

QFITS = os.path.join(os.path.dirname(__file__), "golden", "Q.fits")


def test_reference_qfits_reads_and_rewrites_byte_for_byte(tmp_path):
    hdus = io.read_fits(QFITS)
    assert [h[0] for h in hdus] == [None, "BZ"]
    q, bz = hdus[0][1], hdus[1][1]
    assert q.shape == (100, 240) and q.dtype == np.float32 and bz.shape == (100, 240)
    # the file's documented distribution (median 5.09, min 1.86, max 6.19e6)
    assert abs(float(np.median(q)) - 5.09) < 0.01 and 1.86 < q.min() < 1.87 and 6.1e6 < q.max() < 6.2e6
    out = tmp_path / "Q.fits"
    io.write_fits(out, q, [("Bz", bz)])
    assert out.read_bytes() == open(QFITS, "rb").read()


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int16, np.uint8, np.int64])
def test_fits_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(1)
    a = (rng.normal(size=(3, 5, 7)) * 50).astype(dtype)
    b = np.asfortranarray(rng.normal(size=(4, 9)).astype(np.float64))
    p = tmp_path / "x.fits"
    io.write_fits(p, a, {"Bz": b, "twist": b[:2]})
    hdus = io.read_fits(p)
    assert [h[0] for h in hdus] == [None, "BZ", "TWIST"]
    np.testing.assert_array_equal(hdus[0][1], a)
    np.testing.assert_array_equal(hdus[1][1], b)
    np.testing.assert_array_equal(hdus[2][1], b[:2])
    assert os.path.getsize(p) % 2880 == 0


def test_npz_staggered(tmp_path):
    rng = np.random.default_rng(2)
    nx, ny, nz = 6, 5, 4
    Bx, By, Bz = rng.normal(size=(nx + 1, ny, nz)), rng.normal(size=(nx, ny + 1, nz)), rng.normal(size=(nx, ny, nz + 1))
    p = tmp_path / "f.npz"
    np.savez(p, x=np.arange(nx), y=np.arange(ny), z=np.arange(nz), Bx=Bx, By=By, Bz=Bz, num_ghost_cells=2)
    (bx, by, bz), (x, y, z), ghost = io.read_npz(p, scale=50.0)
    np.testing.assert_array_equal(bx, (Bx[1:].astype(np.float32) * np.float32(50)))
    np.testing.assert_array_equal(by, (By[:, 1:].astype(np.float32) * np.float32(50)))
    np.testing.assert_array_equal(bz, (Bz[:, :, 1:].astype(np.float32) * np.float32(50)))
    assert bx.flags.f_contiguous and bx.dtype == np.float32 and int(ghost) == 2
    assert x.dtype == np.float32 and z.shape == (nz,)


def _idl_cube(rng, shape):
    return rng.normal(size=shape[::-1]).astype(np.float64)  # IDL order: (nz, ny, nx)


def test_read_sav_transposes(monkeypatch):
    import scipy.io as sio
    rng = np.random.default_rng(3)
    shape = (5, 4, 3)
    v = {k: _idl_cube(rng, shape) for k in ("bx", "by", "bz")}
    v.update(xa=np.linspace(0, 1, 5), ya=np.linspace(0, 2, 4), za=np.linspace(0, 3, 3))
    monkeypatch.setattr(sio, "readsav", lambda path: dict(v))
    (bx, by, bz), axes = io.read_sav("any.sav")
    assert bx.shape == shape and bx.dtype == np.float32 and bx.flags.f_contiguous
    np.testing.assert_array_equal(bx, np.transpose(v["bx"], (2, 1, 0)).astype(np.float32))
    np.testing.assert_array_equal(axes[2], v["za"].astype(np.float32))
    # a structure array entry (twbox[0]): upper-case fields, lower-case titles
    dt = np.dtype([(("bx", "BX"), "O"), (("by", "BY"), "O"), (("bz", "BZ"), "O")])
    rec = np.rec.array([(v["bx"], v["by"], v["bz"])], dtype=dt)
    monkeypatch.setattr(sio, "readsav", lambda path: {"twbox": rec})
    (bx2, by2, bz2), axes2 = io.read_sav("any.sav", struct="twbox")
    assert axes2 is None
    np.testing.assert_array_equal(bz2, np.transpose(v["bz"], (2, 1, 0)).astype(np.float32))


def test_save_qube(tmp_path):
    q = np.ones((3, 4, 5), np.float32)
    p = tmp_path / "q.npz"
    io.save_qube(p, q, Twube=q * 2, xi=np.arange(3), B0z=q[:, :, 0])
    with np.load(p) as d:
        assert set(d.files) == {"Qube", "Twube", "xi", "B0z"}
        np.testing.assert_array_equal(d["Twube"], q * 2)
