"""Minimal, independent MATLAB v5 (uncompressed) reader used to cross-check the
library's C++ reader on the reference fixtures (scipy refuses some of them)."""
from __future__ import annotations

import struct

import numpy as np

_DT = {1: np.int8, 2: np.uint8, 3: np.int16, 4: np.uint16, 5: np.int32, 6: np.uint32, 7: np.float32,
       9: np.float64, 12: np.int64, 13: np.uint64}


def _elements(buf: bytes, off: int, end: int):
    while off + 8 <= end:
        t, n = struct.unpack_from("<II", buf, off)
        if t >> 16:  # small data element: size in the upper half, data in the next 4 bytes
            n, t = t >> 16, t & 0xFFFF
            yield t, buf[off + 4: off + 4 + n]
            off += 8
        else:
            yield t, buf[off + 8: off + 8 + n]
            off += 8 + n + ((8 - n % 8) % 8)


def read_mat_v5(path):
    """Returns ((nrows, ncols), rows, cols, vals) for sparse, or (shape, array) for dense."""
    buf = open(path, "rb").read()
    t, n = struct.unpack_from("<II", buf, 128)
    assert t == 14, "not an miMATRIX element"
    subs = list(_elements(buf, 136, 136 + n))
    flags = np.frombuffer(subs[0][1], dtype=np.uint32)
    mclass = int(flags[0]) & 0xFF
    dims = np.frombuffer(subs[1][1], dtype=_DT[subs[1][0]])
    if mclass == 5:  # sparse: ir, jc, pr
        ir = np.frombuffer(subs[3][1], dtype=_DT[subs[3][0]]).astype(np.int64)
        jc = np.frombuffer(subs[4][1], dtype=_DT[subs[4][0]]).astype(np.int64)
        pr = np.frombuffer(subs[5][1], dtype=_DT[subs[5][0]]).astype(np.float64)
        nr, nc = int(dims[0]), int(dims[1])
        nnz = int(jc[nc])
        cols = np.repeat(np.arange(nc), np.diff(jc[:nc + 1]))
        return (nr, nc), ir[:nnz], cols, pr[:nnz]
    data = np.frombuffer(subs[3][1], dtype=_DT[subs[3][0]]).astype(np.float64)
    return tuple(int(d) for d in dims), data
