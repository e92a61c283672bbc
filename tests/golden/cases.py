"""Seeded inputs of the small golden cases (shared by make_golden.py and the tests)."""
from __future__ import annotations

import numpy as np

from vectorcache_amd import synthetic

# (bits, dim, n, with TQ+ calibration)
CASES = [(4, 128, 300, False), (3, 128, 300, False), (2, 128, 300, False), (4, 200, 250, False),
         (4, 768, 200, True), (2, 768, 200, True), (3, 1536, 150, True), (4, 1536, 150, True),
         (2, 1536, 150, False), (4, 8, 64, False), (2, 1536, 150, True)]
KS = [1, 10, 64, 1000]


def case_data(bits: int, dim: int, n: int, idx: int):
    """Database, queries and a calibration sample with the edge cases the tests need."""
    x = synthetic.mixture(n, dim, 16, 0.3, 7000 + idx, 7100 + idx)
    x[3] = 0.0                                  # zero vector: alpha = 0
    x[7] = x[11]                                # exact duplicate: score tie
    x[9] = x[12] * np.float32(1e-12)            # below the 1e-10 norm floor
    q = synthetic.mixture(6, dim, 16, 0.3, 7000 + idx, 7200 + idx)
    q[2] = x[11]                                # query equal to a duplicated vector
    q[4] = 0.0                                  # zero query
    sample = synthetic.mixture(400, dim, 16, 0.3, 7000 + idx, 7300 + idx)
    return x, q, sample


RESCORE_DIMS = (8, 24, 128, 1536)


# score_prepared with every argument: (bits, dim, n, nq, k)
SCORE_PREPARED_CASES = ((4, 128, 3000, 5, 10), (2, 256, 1000, 4, 64), (3, 128, 777, 3, 200))


def score_prepared_inputs(bits, dim, n, nq):
    """Seeded vectors, queries, an id map, a query subset with a repeat, and
    residual norms / scales (the landmark path's arguments)."""
    x = synthetic.mixture(n, dim, 16, 0.3, 6300 + n, 6400 + dim)
    q = synthetic.mixture(nq, dim, 16, 0.3, 6300 + n, 6500 + nq)
    r = np.random.default_rng(6600 + n)
    id_map = (r.permutation(n).astype(np.uint64) * 7 + 3).astype(np.uint64)
    query_indices = np.array([nq - 1, 0, nq - 1], np.uint64)
    residual_norms = (r.random(n) * 0.5).astype(np.float32)
    residual_scales = (r.random(nq) * 2).astype(np.float32)
    return x, q, id_map, query_indices, residual_norms, residual_scales


def SCORE_PREPARED_VARIANTS(id_map, qi, rn, rs):
    """(tag, keyword arguments) of the score_prepared golden outputs."""
    return (("plain", {}), ("idmap", dict(id_map=id_map)), ("subset", dict(query_indices=qi)),
            ("resid", dict(residual_norms=rn, residual_scales=rs)),
            ("all", dict(id_map=id_map, query_indices=qi, residual_norms=rn, residual_scales=rs)),
            ("norms_only", dict(residual_norms=rn)))


def rescore_inputs(dim):
    """Seeded float_rescore_topk inputs: corpus (300 x dim), 6 queries, rows of
    12 candidates with ids past the corpus and a repeated id, per-row counts."""
    r = np.random.default_rng(6100 + dim)
    corpus = r.standard_normal((300, dim)).astype(np.float32)
    q = r.standard_normal((6, dim)).astype(np.float32)
    ids = r.integers(0, 320, (6, 12)).astype(np.uint64)
    ids[1, 3] = ids[1, 4]
    scores = r.standard_normal((6, 12)).astype(np.float32)
    counts = np.array([12, 5, 0, 12, 7, 1], np.uint64)
    return corpus, q, scores, ids, counts


# Landmark index cases: (tag, n, nq, dim, n_landmarks, scan_bits, rerank_bits,
# seed, k, batches, prepare_between, float corpus)
LANDMARK_CASES = (
    ("m20", 500, 7, 64, 20, 4, 0, 11, 10, None, False, False),
    ("m100", 700, 5, 128, 100, 4, 4, 3, 25, None, False, False),
    ("b2", 400, 4, 96, 96, 2, 3, 0, 7, None, False, False),
    ("b3k", 300, 3, 64, 13, 3, 2, 5, 300, None, False, False),
    ("batched", 600, 6, 256, 64, 4, 4, 9, 10, (200, 250, 150), False, False),
    ("reprep", 600, 6, 256, 64, 4, 4, 9, 10, (200, 250, 150), True, False),
    ("rescore", 300, 4, 64, 36, 4, 4, 1, 5, None, False, True),
    ("d1536", 1500, 8, 1536, 1024, 4, 4, 0, 10, None, False, False),
    ("d768b3", 1200, 5, 768, 256, 3, 2, 4, 64, None, False, False),
    ("d384b2", 900, 4, 384, 200, 2, 4, 6, 100, (300, 600), False, False),
    ("m47", 400, 5, 96, 47, 4, 0, 13, 10, None, False, False),
    ("m17b3", 300, 4, 64, 17, 3, 4, 2, 12, None, False, False),
)


# The staged sign-plane branch (the reference on an AVX-512 VNNI host):
# (tag, n, nq, dim, n_landmarks, seed, k); n >= 32768 so the index keeps sign
# planes, and n_landmarks % 32 == 0, the scan widths the reference's sign
# kernels cover.
STAGED_CASES = (
    ("s128k10", 40000, 6, 128, 64, 3, 10),
    ("s128k50", 40000, 6, 128, 64, 3, 50),
    ("s192k100", 36000, 5, 256, 192, 8, 100),
    ("s64k300", 33000, 4, 64, 64, 1, 300),
)

# Scan widths that are not a multiple of 32 (the reference's sign scan reads
# past its sign codes there): the planes index must give the exact scan's result.
UNSTAGED_PLANES_CASE = ("s200k100", 36000, 5, 256, 200, 8, 100)


def landmark_inputs(n, nq, dim):
    """Seeded vectors (row scales varied) and queries with the structured rows
    the reference's quantizers tie on: a zero row and rows whose rotation has
    exact zeros (constant and two-level rows), some repeated as queries."""
    r = np.random.default_rng(7100 + n + dim)
    x = synthetic.mixture(n, dim, 16, 0.3, 7200 + n, 7300 + dim)
    x *= r.uniform(0.5, 2.0, (n, 1)).astype(np.float32)
    q = synthetic.mixture(nq, dim, 16, 0.3, 7200 + n, 7400 + nq)
    x[0] = 1.0
    x[1, ::2] = 2.0
    x[1, 1::2] = 0.0
    x[2, : dim // 2] = -1.0
    x[2, dim // 2:] = 0.5
    x[3] = 0.0
    q[0] = x[0]
    if nq > 2:
        q[2] = x[2]
    return np.ascontiguousarray(x, np.float32), np.ascontiguousarray(q, np.float32)
