import hashlib

import numpy as np


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def digest(a) -> np.ndarray:
    return np.frombuffer(bytes.fromhex(sha(a)), np.uint8)


def bits_equal(a, b) -> bool:
    """Bit-for-bit equality of two float arrays (shape and dtype included)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def scores_equal(a, b) -> bool:
    """Float equality where -0.0 == +0.0 (the library reports -0.0 as +0.0);
    every other score must match bit for bit."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    if a.shape != b.shape:
        return False
    za, zb = a == 0, b == 0
    return bool(np.array_equal(za, zb) and np.array_equal(a[~za].view(np.uint32), b[~zb].view(np.uint32)))
