"""Tagged binary state dumps.

One record = 16-byte space-padded name, int64 kind (1 = float64, 2 = int64),
int64 count, then ``count`` little-endian values.  The GPU host path, the C
oracle and the reference-linked golden driver all write this format so their
results can be compared name by name.
"""
from __future__ import annotations

import numpy as np


def read_dump(path: str) -> dict[str, np.ndarray]:
    raw = open(path, "rb").read()
    out: dict[str, np.ndarray] = {}
    pos = 0
    while pos < len(raw):
        name = raw[pos:pos + 16].decode("ascii").strip()
        kind, count = np.frombuffer(raw, dtype="<i8", count=2, offset=pos + 16)
        pos += 32
        dt = "<f8" if kind == 1 else "<i8"
        out[name] = np.frombuffer(raw, dtype=dt, count=int(count), offset=pos).copy()
        pos += 8 * int(count)
    return out


def write_dump(path: str, records: dict) -> None:
    with open(path, "wb") as f:
        for name, val in records.items():
            a = np.atleast_1d(np.asarray(val))
            if a.dtype.kind == "f":
                kind, a = 1, a.astype("<f8")
            else:
                kind, a = 2, a.astype("<i8")
            f.write(name.ljust(16)[:16].encode("ascii"))
            f.write(np.array([kind, a.size], dtype="<i8").tobytes())
            f.write(np.ascontiguousarray(a.ravel(order="F")).tobytes())
