"""``.mpmbin`` point-cloud files (the reference's checkpoint/interchange format).

The reference writes a PointCloud with cereal's binary archive
(PointCloud.cpp:334-362): a little-endian uint64 point count, then per point
the fields of MaterialPoint::serialize (MaterialPoint.h:52-74) in order, with
Eigen vectors as (int32 rows, floats) and matrices as (int32 rows, int32 cols,
floats row by row) (cereal_eigen.h:10-38).  432 bytes per point.
"""
from __future__ import annotations

import numpy as np

_V = [("r", "<i4"), ("d", "<f4", (3,))]
_M = [("r", "<i4", (2,)), ("d", "<f4", (9,))]

RECORD = np.dtype([
    ("x", _V), ("v", _V), ("F", _M), ("C", _M), ("P", _M),
    ("m", "<f4"), ("vol", "<f4"), ("lam", "<f4"), ("mu", "<f4"),
    ("dFc", _M), ("dLdx", _V), ("dLdv", _V), ("dLdF", _M),
    ("momentum", _M), ("vector", _M), ("vector_max", _M),
])
assert RECORD.itemsize == 432

VEC_FIELDS = ("x", "v", "dLdx", "dLdv")
MAT_FIELDS = ("F", "C", "P", "dFc", "dLdF", "momentum", "vector", "vector_max")
SCALAR_FIELDS = ("m", "vol", "lam", "mu")


def write(path, n=None, **fields):
    """Write a point cloud; missing fields are zero (F defaults to I)."""
    if n is None:
        n = len(next(iter(fields.values())))
    rec = np.zeros(n, dtype=RECORD)
    for f in VEC_FIELDS:
        rec[f]["r"] = 3
        if f in fields:
            rec[f]["d"] = np.asarray(fields[f], dtype=np.float32).reshape(n, 3)
    for f in MAT_FIELDS:
        rec[f]["r"] = 3
        if f in fields:
            rec[f]["d"] = np.asarray(fields[f], dtype=np.float32).reshape(n, 9)
        elif f == "F":
            rec[f]["d"] = np.tile(np.eye(3, dtype=np.float32).reshape(9), (n, 1))
    for f in SCALAR_FIELDS:
        if f in fields:
            rec[f] = np.asarray(fields[f], dtype=np.float32).reshape(n)
    with open(path, "wb") as fh:
        fh.write(np.uint64(n).astype("<u8").tobytes())
        fh.write(rec.tobytes())


def read(path):
    """Read a point cloud into a dict of float64 arrays (vectors (n,3), matrices (n,3,3))."""
    with open(path, "rb") as fh:
        raw = fh.read()
    n = int(np.frombuffer(raw[:8], dtype="<u8")[0])
    if len(raw) != 8 + n * RECORD.itemsize:
        raise ValueError(f"{path}: {len(raw)} bytes for {n} points")
    rec = np.frombuffer(raw[8:], dtype=RECORD, count=n)
    if n and (np.any(rec["x"]["r"] != 3) or np.any(rec["F"]["r"] != 3)):
        raise ValueError(f"{path}: unexpected Eigen shapes")
    out = {}
    for f in VEC_FIELDS:
        out[f] = rec[f]["d"].astype(np.float64)
    for f in MAT_FIELDS:
        out[f] = rec[f]["d"].astype(np.float64).reshape(n, 3, 3)
    for f in SCALAR_FIELDS:
        out[f] = rec[f].astype(np.float64)
    return out
