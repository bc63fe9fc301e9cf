"""NIfTI-1 (.nii) read/write with numpy, for preparing CLI inputs and reading its outputs.

Volumes are returned as (Z, Y, X) arrays so that ``a.ravel()`` is x-fastest, the solver's layout.
"""
from __future__ import annotations

import struct

import numpy as np

_DT = {np.dtype(np.uint8): 2, np.dtype(np.int16): 4, np.dtype(np.int32): 8, np.dtype(np.float32): 16,
       np.dtype(np.float64): 64, np.dtype(np.int8): 256, np.dtype(np.uint16): 512, np.dtype(np.uint32): 768}
_DT_INV = {v: k for k, v in _DT.items()}


def write(path, vol, pixdim=(1.0, 1.0, 1.0)):
    """Write a (Z, Y, X) array as a single-file NIfTI-1 image (little endian, vox_offset 352)."""
    vol = np.ascontiguousarray(vol)
    if vol.dtype not in _DT:
        raise ValueError(f"unsupported dtype {vol.dtype}")
    z, y, x = vol.shape
    hdr = bytearray(348)
    struct.pack_into("<i", hdr, 0, 348)
    hdr[38] = ord("r")
    struct.pack_into("<8h", hdr, 40, 3, x, y, z, 1, 1, 1, 1)
    struct.pack_into("<hh", hdr, 70, _DT[vol.dtype], vol.dtype.itemsize * 8)
    struct.pack_into("<8f", hdr, 76, 1.0, *pixdim, 1.0, 1.0, 1.0, 1.0)
    struct.pack_into("<f", hdr, 108, 352.0)
    struct.pack_into("<ff", hdr, 112, 1.0, 0.0)
    hdr[123] = 2   # mm
    hdr[344:348] = b"n+1\x00"
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(b"\x00" * 4)
        f.write(vol.astype(vol.dtype.newbyteorder("<"), copy=False).tobytes())


def read(path):
    """Read a single-file NIfTI-1 image as a (Z, Y, X) array of its stored dtype."""
    with open(path, "rb") as f:
        raw = f.read()
    (size,) = struct.unpack_from("<i", raw, 0)
    end = "<" if size == 348 else ">"
    dims = struct.unpack_from(end + "8h", raw, 40)
    dt, _ = struct.unpack_from(end + "hh", raw, 70)
    (off,) = struct.unpack_from(end + "f", raw, 108)
    x, y, z = dims[1], dims[2], dims[3]
    dtype = _DT_INV[dt].newbyteorder(end)
    a = np.frombuffer(raw, dtype=dtype, count=x * y * z, offset=int(off))
    return a.reshape(z, y, x).astype(dtype.newbyteorder("="))
