"""Helpers shared by the fixture generator and the tests."""
import hashlib

import numpy as np


def edge_checksum(src, dst) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(src, np.int32).tobytes())
    h.update(np.ascontiguousarray(dst, np.int32).tobytes())
    return h.hexdigest()[:16]


def rel_err(a, b) -> float:
    return abs(a - b) / max(abs(b), 1e-300) if b != 0 else abs(a)
