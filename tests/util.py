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


def score_into_scenarios(golden):
    """The chained score_prepared_into scenarios of tests/golden/score_into.npz:
    yields (name, bits, dim, k, q, shift, scale, calls, want) where calls is a
    list of (perm0, alpha, id_map, query_indices, offsets or None) and want[ci]
    the reference's five heap arrays after call ci, plus want["final"] =
    heaps_to_search_results of the last state."""
    names = sorted({key.split("_")[0] for key in golden})
    for name in names:
        q = golden[f"{name}_q"]
        dim = q.shape[1]
        shift = golden.get(f"{name}_shift")
        scale = golden.get(f"{name}_scale")
        calls = []
        ci = 0
        while f"{name}_c{ci}_perm0" in golden:
            calls.append((golden[f"{name}_c{ci}_perm0"], golden[f"{name}_c{ci}_alpha"], golden[f"{name}_c{ci}_idmap"],
                          golden[f"{name}_c{ci}_qidx"], golden.get(f"{name}_c{ci}_offsets")))
            ci += 1
        ks = sorted({int(key.split("_")[1][1:]) for key in golden if key.startswith(f"{name}_k")})
        bits = int(golden[f"{name}_bits"][0])
        for k in ks:
            want = {}
            for c in range(len(calls)):
                want[c] = tuple(golden[f"{name}_k{k}_c{c}_{h}"] for h in ("hs", "hi", "hsz", "hmin", "hmi"))
            want["final"] = (golden[f"{name}_k{k}_scores"], golden[f"{name}_k{k}_ids"])
            yield name, bits, dim, k, q, shift, scale, calls, want


def empty_heaps(nq, k, seeded=None):
    """Fresh reference heaps: all -inf thresholds, seeded = {qi: threshold}."""
    heaps = (np.zeros((nq, k), np.float32), np.zeros((nq, k), np.uint64), np.zeros(nq, np.uint64),
             np.full(nq, -np.inf, np.float32), np.zeros(nq, np.uint64))
    for qi, t in (seeded or {}).items():
        heaps[3][qi] = np.float32(t)
    return heaps


def heaps_equal(a, b) -> bool:
    return (bits_equal(np.asarray(a[0], np.float32), np.asarray(b[0], np.float32))
            and all(np.array_equal(np.asarray(x, np.uint64), np.asarray(y, np.uint64)) for x, y in
                    ((a[1], b[1]), (a[2], b[2]), (a[4], b[4])))
            and bits_equal(np.asarray(a[3], np.float32), np.asarray(b[3], np.float32)))
