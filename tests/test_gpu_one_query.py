"""One-query scans with k <= 64 keep a top-k list per wavefront (no
workgroup admission per step) and merge the workgroup's lists at the end.
Their results must be those of the same queries in one batch (tiles of 8
queries: the workgroup admission) and of the oracle's search_flat
(src/search/search.cpp:1293-1302) bit for bit, ties included: ranks are
(score desc, id asc).

Shapes: 128-d, 768-d and 1536-d 4-bit, 1536-d 2-bit, an index of 70 distinct
rows repeated (every score tied 1000 times), k of 1, 10, 64 (the largest
with wavefront lists) and 65 (the workgroup admission), and sizes that leave
a ragged last 64-vector tile.
"""
import numpy as np
import pytest

from util import scores_equal

pytestmark = pytest.mark.gpu


def _rows(n, dim, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, dim)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x


def _one_at_a_time(ix, q, k):
    out_s, out_i = [], []
    for j in range(q.shape[0]):
        r = ix.search(q[j: j + 1], k)
        out_s.append(r.scores[0])
        out_i.append(r.ids[0])
    return np.stack(out_s), np.stack(out_i)


@pytest.mark.parametrize("n,dim,bits", [(100_003, 128, 4), (65_536, 768, 4), (70_001, 1536, 4), (80_000, 1536, 2)])
def test_one_query_scan_equals_batch(vc, n, dim, bits):
    x = _rows(n, dim, 11 + dim + bits)
    q = _rows(5, dim, 99)
    q[2] = x[n - 1]  # the last vector, in the ragged tile
    ix = vc.TurboQuantIndex(dim, bits)
    ix.add(x)
    ix.prepare()
    vc.set_profiling(True)
    for k in (1, 10, 64, 65):
        s1, i1 = _one_at_a_time(ix, q, k)
        _, tile, ranges = vc.last_scan_stats()
        assert tile == 1 and ranges >= 1
        batch = ix.search(q, k)  # a tile of 8 queries: the workgroup admission
        assert np.array_equal(batch.ids, i1) and scores_equal(batch.scores, s1), (n, dim, bits, k)
    vc.set_profiling(False)
    ix.close()


def test_one_query_scan_matches_oracle(vc, oracle):
    n, dim, bits = 100_003, 128, 4
    x = _rows(n, dim, 5)
    q = _rows(4, dim, 6)
    ix = vc.TurboQuantIndex(dim, bits)
    ix.add(x)
    ix.prepare()
    perm0, alpha = ix.export_perm0(), ix.export_alpha()
    for k in (1, 10, 64, 65):
        s1, i1 = _one_at_a_time(ix, q, k)
        s, i = oracle.search_flat(q, k, bits, perm0, alpha)
        assert np.array_equal(i1, i) and scores_equal(s1, s), k
    ix.close()


def test_one_query_scan_ties(vc, oracle):
    """70 distinct rows, each repeated 1000 times: every score is tied 1000
    ways, so the top k is decided by the lowest ids."""
    dim, bits = 128, 4
    base = _rows(70, dim, 7)
    x = np.repeat(base, 1000, axis=0)
    rng = np.random.default_rng(8)
    x = x[rng.permutation(x.shape[0])]
    q = np.concatenate([base[:3], _rows(2, dim, 9)])
    ix = vc.TurboQuantIndex(dim, bits)
    ix.add(x)
    ix.prepare()
    perm0, alpha = ix.export_perm0(), ix.export_alpha()
    for k in (1, 10, 64, 65):
        s1, i1 = _one_at_a_time(ix, q, k)
        batch = ix.search(q, k)
        assert np.array_equal(batch.ids, i1) and scores_equal(batch.scores, s1), k
        s, i = oracle.search_flat(q, k, bits, perm0, alpha)
        assert np.array_equal(i1, i) and scores_equal(s1, s), k
    ix.close()
