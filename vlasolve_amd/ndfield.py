"""NDfield binary snapshots (the reference's snapshot format).

Layout as written by Save_NDfield (NDfield.hxx:308-362; manual/FILE_FORMAT):
a Fortran-style block holding the 16-byte tag "NDFIELD", a 652-byte header
block (comment[80], ndims, dims[20], fdims_index, datatype, x0[20],
delta[20], 160 spare bytes) and the data block, each block framed by its
byte count.
"""
from __future__ import annotations

import struct

import numpy as np

MAX_DIMS = 20
ND_DOUBLE = 1 << 9
_DTYPES = {1: np.int8, 2: np.uint8, 4: np.int16, 8: np.uint16, 16: np.int32, 32: np.uint32,
           64: np.int64, 128: np.uint64, 256: np.float32, 512: np.float64}
HEADER_BYTES = 4 * (MAX_DIMS + 3) + 8 * (2 * MAX_DIMS) + 160 + 80


def read(path):
    """Return (array, meta).  The array is in file order with dims[0] fastest,
    i.e. shape (dims[ndims-1], ..., dims[0])."""
    with open(path, "rb") as fh:
        raw = fh.read()
    o = 0
    (n,) = struct.unpack_from("<i", raw, o)
    o += 4
    tag = raw[o:o + 16].split(b"\0")[0].decode()
    o += 16 + 4
    if tag != "NDFIELD" or n != 16:
        raise ValueError(f"{path}: not an NDfield file")
    (hb,) = struct.unpack_from("<i", raw, o)
    o += 4
    comment = raw[o:o + 80].split(b"\0")[0].decode(errors="replace")
    o += 80
    (ndims,) = struct.unpack_from("<i", raw, o)
    o += 4
    dims = struct.unpack_from(f"<{MAX_DIMS}i", raw, o)[:ndims]
    o += 4 * MAX_DIMS
    fdims_index, datatype = struct.unpack_from("<ii", raw, o)
    o += 8
    x0 = struct.unpack_from(f"<{MAX_DIMS}d", raw, o)[:ndims]
    o += 8 * MAX_DIMS
    delta = struct.unpack_from(f"<{MAX_DIMS}d", raw, o)[:ndims]
    o += 8 * MAX_DIMS + 160 + 4
    (nbytes,) = struct.unpack_from("<i", raw, o)
    o += 4
    dt = np.dtype(_DTYPES[datatype])
    count = int(np.prod(dims))
    arr = np.frombuffer(raw, dtype=dt, count=count, offset=o).copy()
    meta = dict(comment=comment, dims=list(dims), x0=list(x0), delta=list(delta),
                fdims_index=fdims_index, datatype=datatype, header_bytes=hb, data_bytes=nbytes)
    return arr.reshape(tuple(reversed(dims))), meta


def write(path, arr, x0, delta, comment=""):
    """Write a float64 array (shape reversed dims, dims[0] fastest)."""
    a = np.ascontiguousarray(arr, np.float64)
    dims = list(reversed(a.shape))
    ndims = len(dims)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<i", 16))
        fh.write(b"NDFIELD".ljust(16, b"\0"))
        fh.write(struct.pack("<i", 16))
        fh.write(struct.pack("<i", HEADER_BYTES))
        fh.write(comment.encode()[:79].ljust(80, b"\0"))
        fh.write(struct.pack("<i", ndims))
        fh.write(struct.pack(f"<{MAX_DIMS}i", *(dims + [0] * (MAX_DIMS - ndims))))
        fh.write(struct.pack("<ii", 0, ND_DOUBLE))
        fh.write(struct.pack(f"<{MAX_DIMS}d", *(list(x0) + [0.0] * (MAX_DIMS - ndims))))
        fh.write(struct.pack(f"<{MAX_DIMS}d", *(list(delta) + [0.0] * (MAX_DIMS - ndims))))
        fh.write(b"\0" * 160)
        fh.write(struct.pack("<i", HEADER_BYTES))
        nb = a.size * 8
        fh.write(struct.pack("<i", nb & 0x7FFFFFFF if nb >= 2**31 else nb))
        fh.write(a.tobytes())
        fh.write(struct.pack("<i", nb & 0x7FFFFFFF if nb >= 2**31 else nb))
