"""Field ingest and result files in the formats the reference's notebooks use.

* ``read_sav``   IDL save files through ``scipy.io.readsav`` with the
  ``np.transpose(a, (2, 1, 0))`` that gives (nx, ny, nz) arrays with x
  fastest in memory (demo/Q_RKF45CuPy.ipynb cells [1]-[3]; the stretched
  grid's ``xa``/``ya``/``za`` in demo/Q_RKF45CuPyStretch.ipynb cells [2]-[3];
  the ``twbox[0]`` structure of demo/Q_RKF45MixCuPyDiffScott.ipynb cell [3]).
* ``read_npz``   a NumPy archive with x, y, z and the staggered components
  Bx, By, Bz trimmed to the nodes (demo/npzinputRKF45CuPyTwist.ipynb cell [3]).
* ``write_fits`` / ``read_fits``  FITS images: the Q map as the primary HDU
  and named image extensions (demo/Q_RKF45CuPyScott.ipynb cell [14] writes
  q_0 with a ``Bz`` extension).  The writer produces the same bytes as
  astropy's PrimaryHDU / ImageHDU for these arrays (astropy is not needed).
* ``save_qube``  the 3-D result archive of demo/npzinputRKF45CuPyTwist.ipynb
  cell [30] (``np.savez`` of Qube, Twube and the axes).

These run on the host: they move files, not field lines.
"""
from __future__ import annotations

import numpy as np

_BITPIX = {np.dtype(np.uint8): 8, np.dtype(np.int16): 16, np.dtype(np.int32): 32, np.dtype(np.int64): 64,
           np.dtype(np.float32): -32, np.dtype(np.float64): -64}
_DTYPE = {v: k for k, v in _BITPIX.items()}
_BLOCK = 2880


def _member(src, key):
    """src[key] for readsav's dict of variables or a record of a structure
    array (whose fields are upper case, with the lower-case names as titles)."""
    for k in (key, key.upper(), key.lower()):
        try:
            return src[k]
        except (KeyError, ValueError, IndexError, TypeError):
            pass
    return getattr(src, key)


def read_sav(path, struct=None, index=0, keys=("bx", "by", "bz"), axes_keys=("xa", "ya", "za")):
    """(Bx, By, Bz) as float32 (nx, ny, nz) arrays (x fastest in memory) from
    an IDL save file, and the axes (float32) when the file holds them (else
    None).  ``struct``/``index`` select a structure array entry such as
    ``twbox[0]``."""
    import scipy.io as sio
    a = sio.readsav(path)
    src = _member(a, struct)[index] if struct is not None else a
    field = tuple(np.transpose(np.asarray(_member(src, k)), (2, 1, 0)).astype(np.float32) for k in keys)
    try:
        axes = tuple(np.asarray(_member(src, k)).astype(np.float32) for k in axes_keys)
    except (KeyError, AttributeError):
        axes = None
    return field, axes


def read_npz(path, scale=1.0, staggered=True):
    """(Bx, By, Bz) float32 (nx, ny, nz), Fortran-ordered, and the node
    coordinates (x, y, z) from a NumPy archive.  With ``staggered`` the
    components come on face-centred grids one longer along their own axis
    and the first plane is dropped (Bx[1:], By[:, 1:], Bz[:, :, 1:]), as the
    notebook does; ``scale`` multiplies the field.  Also returns the
    archive's ``num_ghost_cells`` (None if absent)."""
    with np.load(path, allow_pickle=False) as d:
        x, y, z = (d[k].astype(np.float32) for k in ("x", "y", "z"))
        bx, by, bz = d["Bx"], d["By"], d["Bz"]
        if staggered:
            bx, by, bz = bx[1:, :, :], by[:, 1:, :], bz[:, :, 1:]
        field = tuple(np.asfortranarray(b.astype(np.float32) * np.float32(scale) if scale != 1.0
                                        else b.astype(np.float32)) for b in (bx, by, bz))
        ghost = d["num_ghost_cells"] if "num_ghost_cells" in d.files else None
    if len({b.shape for b in field}) != 1:
        raise ValueError(f"components have different shapes {[b.shape for b in field]}")
    return field, (x, y, z), ghost


def _card(key, value=None, comment=None):
    if value is None:
        s = key.ljust(8)
    else:
        if isinstance(value, bool):
            v = ("T" if value else "F").rjust(20)
        elif isinstance(value, str):
            v = ("'" + value.ljust(8) + "'").ljust(20)
        else:
            v = str(int(value)).rjust(20)
        s = f"{key:<8}= {v}" + (f" / {comment}" if comment else "")
    if len(s) > 80:
        raise ValueError(f"FITS card too long: {s!r}")
    return s.ljust(80)


def _hdu(cards, data):
    head = "".join(cards) + _card("END")
    head = head.ljust(-(-len(head) // _BLOCK) * _BLOCK)
    raw = np.ascontiguousarray(data).astype(data.dtype.newbyteorder(">"), copy=False).tobytes()
    return head.encode("ascii") + raw + b"\0" * (-len(raw) % _BLOCK)


def _image_cards(arr):
    bitpix = _BITPIX.get(arr.dtype.newbyteorder("=")) if arr.dtype.byteorder in "<>" else _BITPIX.get(arr.dtype)
    if bitpix is None:
        raise TypeError(f"no FITS BITPIX for dtype {arr.dtype}")
    cards = [_card("BITPIX", bitpix, "array data type"), _card("NAXIS", arr.ndim, "number of array dimensions")]
    cards += [_card(f"NAXIS{i + 1}", n) for i, n in enumerate(arr.shape[::-1])]
    return cards


def write_fits(path, primary, extensions=(), overwrite=True):
    """A FITS file with ``primary`` as the primary image and each (name,
    array) of ``extensions`` (a sequence or dict) as an IMAGE extension, the
    layout of astropy's HDUList([PrimaryHDU(q), ImageHDU(bz, name='Bz')]):
    C-order data (NAXIS1 = the last axis), big-endian, 2880-byte blocks."""
    import os
    if not overwrite and os.path.exists(path):
        raise FileExistsError(path)
    ext = list(extensions.items()) if isinstance(extensions, dict) else list(extensions)
    p = np.asarray(primary)
    cards = [_card("SIMPLE", True, "conforms to FITS standard")] + _image_cards(p)
    if ext:
        cards.append(_card("EXTEND", True))
    out = [_hdu(cards, p)]
    for name, arr in ext:
        a = np.asarray(arr)
        cards = [_card("XTENSION", "IMAGE", "Image extension")] + _image_cards(a)
        cards += [_card("PCOUNT", 0, "number of parameters"), _card("GCOUNT", 1, "number of groups")]
        if name:
            cards.append(_card("EXTNAME", str(name).upper(), "extension name"))
        out.append(_hdu(cards, a))
    with open(path, "wb") as fh:
        for blob in out:
            fh.write(blob)


def _parse_value(v):
    v = v.split("/", 1)[0].strip() if not v.strip().startswith("'") else v.strip()
    if v.startswith("'"):
        return v[1:v.index("'", 1)].rstrip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        return float(v)


def read_fits(path):
    """[(extname or None, array), ...] of the image HDUs of a FITS file
    (native byte order)."""
    with open(path, "rb") as fh:
        buf = fh.read()
    pos, out = 0, []
    while pos + _BLOCK <= len(buf):
        head = {}
        while True:
            card = buf[pos:pos + 80].decode("ascii")
            pos += 80
            key = card[:8].strip()
            if key == "END":
                break
            if card[8:10] == "= ":
                head[key] = _parse_value(card[10:])
        pos = -(-pos // _BLOCK) * _BLOCK
        shape = tuple(head[f"NAXIS{i}"] for i in range(head.get("NAXIS", 0), 0, -1))
        dt = _DTYPE[head["BITPIX"]].newbyteorder(">")
        nbytes = int(np.prod(shape)) * dt.itemsize if shape else 0
        arr = np.frombuffer(buf, dtype=dt, count=nbytes // dt.itemsize, offset=pos).reshape(shape)
        out.append((head.get("EXTNAME"), arr.astype(dt.newbyteorder("="))))
        pos += -(-nbytes // _BLOCK) * _BLOCK
    return out


def save_qube(path, Qube, Twube=None, xi=None, yi=None, zi=None, **extra):
    """np.savez of a 3-D result: Qube (and Twube, the axes, and any other
    named arrays) under the notebook's keys."""
    arrays = {"Qube": np.asarray(Qube)}
    for k, v in (("Twube", Twube), ("xi", xi), ("yi", yi), ("zi", zi)):
        if v is not None:
            arrays[k] = np.asarray(v)
    arrays.update({k: np.asarray(v) for k, v in extra.items()})
    np.savez(path, **arrays)
