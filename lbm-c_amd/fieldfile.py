"""Reader/writer of the LBMCFLD1 per-node field files used by the host driver
(format in host/field_file.hpp): named float64 / int32 arrays, x fastest."""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"LBMCFLD1"


def write(path, fields: dict):
    """fields: name -> array (float arrays stored as float64, integer arrays as int32)."""
    with open(path, "wb") as fh:
        fh.write(MAGIC)
        fh.write(struct.pack("<I", len(fields)))
        for name, arr in fields.items():
            a = np.asarray(arr).reshape(-1)
            if np.issubdtype(a.dtype, np.integer):
                a, t = a.astype("<i4"), 2
            else:
                a, t = a.astype("<f8"), 1
            nb = name.encode()[:31].ljust(32, b"\0")
            fh.write(nb)
            fh.write(struct.pack("<IQ", t, a.size))
            fh.write(a.tobytes())


def read(path) -> dict:
    out = {}
    with open(path, "rb") as fh:
        if fh.read(8) != MAGIC:
            raise ValueError(f"{path}: not an LBMCFLD1 file")
        (n,) = struct.unpack("<I", fh.read(4))
        for _ in range(n):
            name = fh.read(32).split(b"\0", 1)[0].decode()
            t, count = struct.unpack("<IQ", fh.read(12))
            dt = "<f8" if t == 1 else "<i4"
            out[name] = np.frombuffer(fh.read(count * np.dtype(dt).itemsize), dtype=dt).copy()
    return out
