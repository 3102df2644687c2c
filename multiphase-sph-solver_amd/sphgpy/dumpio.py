"""Reader for the full-precision state dumps ("SPHD" format) written by the oracle harness."""
from __future__ import annotations

import struct
from typing import Dict, Tuple

import numpy as np


def read_dump(path: str) -> Tuple[Dict[str, np.ndarray], int, float]:
    """Return ({field: array (n,) or (n, ncomp)}, step, t)."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:4] != b"SPHD":
        raise ValueError(f"{path}: not an SPHD dump")
    n, step = struct.unpack_from("<ii", buf, 4)
    (t,) = struct.unpack_from("<d", buf, 12)
    (nf,) = struct.unpack_from("<i", buf, 20)
    off = 24
    out: Dict[str, np.ndarray] = {}
    for _ in range(nf):
        name = buf[off:off + 16].split(b"\0", 1)[0].decode()
        ncomp, dt = struct.unpack_from("<ii", buf, off + 16)
        off += 24
        dtype = np.float64 if dt == 0 else np.int32
        cnt = n * ncomp
        arr = np.frombuffer(buf, dtype=dtype, count=cnt, offset=off).copy()
        off += cnt * np.dtype(dtype).itemsize
        out[name] = arr.reshape(n, ncomp) if ncomp > 1 else arr
    return out, step, t
