"""The landmark index (reference LandmarkIndex, src/index.cpp:120-289, on the
AVX2 path): the oracle's restatement against the golden vectors written from
the reference build (tests/golden/landmark.npz, tests/golden/make_golden.py)
and against the mathematics of its basis; the library's host-side basis; and
(marked gpu) the HIP path through the C ABI against both.

Bar: bit-exact scores (-0.0 == +0.0) and ids, landmark basis, 4-bit basis,
query projections, reported sizes.
"""
import numpy as np
import pytest

from cases import LANDMARK_CASES, landmark_inputs
from conftest import load_golden
from util import bits_equal, digest, scores_equal


@pytest.fixture(scope="module")
def golden():
    return load_golden("landmark.npz")


def _lm_kwargs(case):
    tag, n, nq, dim, m, bits, rb, seed, k, batches, pb, fc = case
    return dict(scan_bits=bits, n_landmarks=m, seed=seed, rerank_bits=rb, batches=batches, prepare_between=pb)


# ---- oracle (CPU) ------------------------------------------------------------

@pytest.mark.parametrize("case", LANDMARK_CASES, ids=[c[0] for c in LANDMARK_CASES])
def test_oracle_landmark_search_matches_golden(oracle, golden, case):
    tag, n, nq, dim, m, bits, rb, seed, k, batches, pb, fc = case
    x, q = landmark_inputs(n, nq, dim)
    s, i, info = oracle.lm_search(x, q, k, float_corpus=x if fc else None, **_lm_kwargs(case))
    assert scores_equal(s, golden[f"{tag}_scores"])
    assert np.array_equal(i, golden[f"{tag}_ids"])
    assert [info["scan_dim"], info["codes_bytes"], info["index_bytes"]] == golden[f"{tag}_info"].tolist()


@pytest.mark.parametrize("dim,m,seed", [(64, 20, 0), (128, 100, 7)])
def test_oracle_basis_matches_golden(oracle, golden, dim, m, seed):
    key = f"{dim}_{m}_{seed}"
    L = oracle.ortho_landmarks(dim, m, seed)
    assert np.array_equal(digest(L), golden[f"L_{key}_sha"])
    if f"L_{key}" in golden:
        assert bits_equal(L, golden[f"L_{key}"])
    codes, scales = oracle.quantize_landmarks(L)
    assert np.array_equal(codes, golden[f"Lcodes_{key}"])
    assert bits_equal(scales, golden[f"Lscales_{key}"])
    ql = oracle.project_q4(golden[f"qrot_{key}"], L, (m + 7) // 8 * 8)
    assert bits_equal(ql, golden[f"qland_{key}"])


def test_oracle_basis_1536_matches_golden(oracle, golden):
    assert np.array_equal(digest(oracle.ortho_landmarks(1536, 1024, 0)), golden["L_1536_1024_0_sha"])


@pytest.mark.parametrize("dim,m,seed", [(64, 64, 0), (256, 100, 3)])
def test_oracle_basis_is_orthonormal(oracle, dim, m, seed):
    """Gram-Schmidt in double: the float columns are orthonormal to float
    rounding, and distinct seeds give distinct bases."""
    L = oracle.ortho_landmarks(dim, m, seed).astype(np.float64)
    gram = L @ L.T
    assert np.abs(gram - np.eye(m)).max() < 1e-6
    assert not np.array_equal(L, oracle.ortho_landmarks(dim, m, seed + 1).astype(np.float64))


def test_oracle_projection_against_float64(oracle):
    """The 8-lane float dot of the row projection is within float rounding of
    the exact product, and columns past m are zero."""
    r = np.random.default_rng(3)
    L = oracle.ortho_landmarks(128, 30, 2)
    Y = r.standard_normal((40, 128)).astype(np.float32)
    out = oracle.project_rows(Y, L, 32)
    exact = Y.astype(np.float64) @ L.T.astype(np.float64)
    assert np.abs(out[:, :30] - exact).max() < 1e-4
    assert not out[:, 30:].any()


def test_oracle_q4_projection_closed_form(oracle):
    """With a basis whose columns quantize exactly (levels +-7, 0) and a query
    of integer levels, q_land = (amax / 127) (scale_j) <q_levels, levels_j>."""
    dim, m = 32, 3
    L = np.zeros((m, dim), np.float32)
    L[0, 0], L[0, 5] = 0.5, -0.5
    L[1, :] = 0.25
    L[2, 31] = -1.0
    q = np.zeros((1, dim), np.float32)
    q[0, 0], q[0, 5], q[0, 31], q[0, 7] = 127.0, 64.0, -3.0, 1.0
    ql = oracle.project_q4(q, L, 8)
    lv = np.rint(L / np.abs(L).max(axis=1, keepdims=True) * 7)
    want = (np.float32(127.0 / 127.0) * (np.abs(L).max(axis=1) / np.float32(7.0))) * (lv @ q[0]).astype(np.float32)
    assert np.array_equal(ql[0, :m], want.astype(np.float32))
    assert not ql[0, m:].any()


def test_oracle_matches_reference_build(oracle, ref):
    """Extra seeds and shapes against the reference build itself."""
    r = np.random.default_rng(77)
    # m % 8 in {5, 0, 1, 3, 6, 7}: every tail of the residual norm's 8-wide blocks
    for n, nq, dim, m, bits, rb, k in ((350, 4, 72, 29, 4, 3, 9), (260, 3, 40, 40, 2, 4, 33),
                                       (300, 3, 64, 17, 4, 4, 10), (280, 3, 48, 11, 3, 2, 7),
                                       (320, 4, 96, 30, 4, 3, 12), (300, 3, 56, 47, 2, 4, 15)):
        x = r.standard_normal((n, dim)).astype(np.float32)
        q = r.standard_normal((nq, dim)).astype(np.float32)
        x[5] = 0.5
        a = oracle.lm_search(x, q, k, scan_bits=bits, n_landmarks=m, seed=n, rerank_bits=rb)
        b = ref.lm_search(x, q, k, scan_bits=bits, n_landmarks=m, seed=n, rerank_bits=rb)
        assert scores_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert all(a[2][key] == b[2][key] for key in ("scan_dim", "codes_bytes", "index_bytes"))


# ---- the library's host-side basis (no GPU calls) ----------------------------

def test_library_basis_matches_golden(golden):
    from vectorcache_amd import landmark

    for dim, m, seed in ((64, 20, 0), (128, 100, 7)):
        key = f"{dim}_{m}_{seed}"
        L = landmark.make_ortho_landmarks(dim, m, seed)
        assert np.array_equal(digest(L), golden[f"L_{key}_sha"])
        codes, scales = landmark.quantize_landmarks(L)
        assert np.array_equal(codes, golden[f"Lcodes_{key}"])
        assert bits_equal(scales, golden[f"Lscales_{key}"])
    assert np.array_equal(digest(landmark.make_ortho_landmarks(1536, 1024, 0)), golden["L_1536_1024_0_sha"])
    with pytest.raises(ValueError):
        landmark.make_ortho_landmarks(64, 65, 0)


# ---- HIP path ----------------------------------------------------------------

def _gpu_index(case, x):
    from vectorcache_amd.landmark import LandmarkIndex

    tag, n, nq, dim, m, bits, rb, seed, k, batches, pb, fc = case
    ix = LandmarkIndex(dim, bits, m, seed, rb)
    off = 0
    for b in batches or (n,):
        ix.add(x[off: off + b])
        off += b
        if pb:
            ix.prepare()
    if not pb:
        ix.prepare()
    return ix


@pytest.mark.gpu
@pytest.mark.parametrize("case", LANDMARK_CASES, ids=[c[0] for c in LANDMARK_CASES])
def test_gpu_landmark_search_matches_golden(golden, case):
    tag, n, nq, dim, m, bits, rb, seed, k, batches, pb, fc = case
    print(f"landmark case {tag}", flush=True)
    x, q = landmark_inputs(n, nq, dim)
    ix = _gpu_index(case, x)
    res = ix.search(q, k, float_corpus=x if fc else None)
    assert scores_equal(res.scores, golden[f"{tag}_scores"])
    assert np.array_equal(res.ids, golden[f"{tag}_ids"])
    assert [ix.scan_dim(), ix.codes_bytes(), ix.index_bytes()] == golden[f"{tag}_info"].tolist()


@pytest.mark.gpu
def test_gpu_landmark_device_search_equals_host_search(golden):
    import torch

    case = LANDMARK_CASES[1]
    tag, n, nq, dim, m, bits, rb, seed, k, batches, pb, fc = case
    x, q = landmark_inputs(n, nq, dim)
    ix = _gpu_index(case, torch.from_numpy(x).cuda())
    s, i = ix.search_device(torch.from_numpy(q).cuda(), k)
    torch.cuda.synchronize()
    assert scores_equal(s.cpu().numpy(), golden[f"{tag}_scores"])
    assert np.array_equal(i.cpu().numpy().view(np.uint64), golden[f"{tag}_ids"])
    s2, i2 = ix.search_device(torch.from_numpy(q).cuda(), k, float_corpus=torch.from_numpy(x).cuda())
    h = ix.search(q, k, float_corpus=x)
    torch.cuda.synchronize()
    assert scores_equal(s2.cpu().numpy(), h.scores) and np.array_equal(i2.cpu().numpy().view(np.uint64), h.ids)


@pytest.mark.gpu
def test_gpu_landmark_query_rows_match_oracle(oracle):
    """q_land and q_perp of the query side against the oracle's rotation and
    q4 projection, q_perp by sequential double sums (pure Python)."""
    import math

    from vectorcache_amd.landmark import LandmarkIndex, make_ortho_landmarks

    dim, m, seed = 128, 100, 7
    ix = LandmarkIndex(dim, 4, m, seed)
    _, q = landmark_inputs(50, 6, dim)
    ql, qp = ix.query_rows(q)
    rot = np.stack([oracle.rotate_scaled(r, np.float32(1.0) / np.float32(oracle.row_norm(r))) for r in q])
    want = oracle.project_q4(rot, make_ortho_landmarks(dim, m, seed), ix.scan_dim())
    assert bits_equal(ql, want)
    for j in range(q.shape[0]):
        q2 = 0.0
        for v in rot[j]:
            q2 = q2 + float(v) * float(v)
        l2 = 0.0
        for v in want[j, :m]:
            l2 = l2 + float(v) * float(v)
        assert qp[j] == np.float32(math.sqrt(max(0.0, q2 - l2)))


@pytest.mark.gpu
def test_gpu_landmark_at_1536_matches_oracle(oracle):
    """dim 1536 with the default 1024 landmarks, 20000 vectors, against the oracle."""
    from vectorcache_amd.landmark import LandmarkIndex

    r = np.random.default_rng(8)
    x = r.standard_normal((20000, 1536)).astype(np.float32)
    q = r.standard_normal((6, 1536)).astype(np.float32)
    ix = LandmarkIndex(1536, 4)
    ix.add(x)
    ix.prepare()
    res = ix.search(q, 10)
    s, i, info = oracle.lm_search(x, q, 10, scan_bits=4, n_landmarks=1024, seed=0)
    assert scores_equal(res.scores, s) and np.array_equal(res.ids, i)
    assert ix.index_bytes() == info["index_bytes"]


@pytest.mark.gpu
def test_gpu_landmark_errors():
    from vectorcache_amd._capi import IndexStateError
    from vectorcache_amd.landmark import LandmarkIndex

    for args in ((60, 4), (64, 5), (64, 4, 0), (64, 4, 65), (64, 4, 8, 0, 1)):
        with pytest.raises(ValueError):
            LandmarkIndex(*args)
    ix = LandmarkIndex(64, 4, 20)
    with pytest.raises(ValueError):
        ix.add(np.zeros(65, np.float32))
    ix.add(np.ones((10, 64), np.float32))
    with pytest.raises(IndexStateError):
        ix.search(np.ones((1, 64), np.float32), 3)
    ix.prepare()
    assert ix.search(np.ones((1, 64), np.float32), 0).k == 0
    assert ix.search(np.ones((1, 64), np.float32), 5000).k == 10  # k clipped to size()


@pytest.mark.gpu
def test_gpu_landmark_edge_cases(oracle):
    """Empty index, fewer vectors than k, no queries, and add after prepare
    (the index must be prepared again; earlier rows are zeroed as in the
    reference) against the oracle."""
    from vectorcache_amd._capi import IndexStateError
    from vectorcache_amd.landmark import LandmarkIndex

    ix = LandmarkIndex(64, 4, 20)
    ix.prepare()
    assert ix.search(np.ones((3, 64), np.float32), 10).scores.shape == (3, 0)
    x, q = landmark_inputs(40, 3, 64)
    ix.add(x[:5])
    ix.prepare()
    res = ix.search(q, 10)
    s, i, _ = oracle.lm_search(x[:5], q, 10, n_landmarks=20, seed=0)
    assert res.k == 5 and scores_equal(res.scores, s) and np.array_equal(res.ids, i)
    assert ix.search(np.zeros((0, 64), np.float32), 4).scores.shape == (0, 4)
    ix.add(x[5:40])
    with pytest.raises(IndexStateError):
        ix.search(q, 3)
    ix.prepare()
    res = ix.search(q, 7)
    s, i, _ = oracle.lm_search(x, q, 7, n_landmarks=20, seed=0, batches=(5, 35), prepare_between=True)
    assert scores_equal(res.scores, s) and np.array_equal(res.ids, i)
    assert ix.size() == 40


@pytest.mark.gpu
def test_gpu_landmark_query_rotation_paths_agree(golden):
    """The query rotation inside the projection kernel (default) and in its own
    kernel (debug bit 5) give the same query rows and the golden results."""
    import vectorcache_amd.index as vci

    case = LANDMARK_CASES[-1]  # dim 1536, 1024 landmarks: rotation blocks of 512
    tag, n, nq, dim, m, bits, rb, seed, k, batches, pb, fc = case
    x, q = landmark_inputs(n, nq, dim)
    ix = _gpu_index(case, x)
    fused = ix.query_rows(q)
    res = ix.search(q, k)
    vci.set_debug(32)
    try:
        apart = ix.query_rows(q)
        res2 = ix.search(q, k)
    finally:
        vci.set_debug(0)
    assert bits_equal(fused[0], apart[0]) and bits_equal(fused[1], apart[1])
    for r in (res, res2):
        assert scores_equal(r.scores, golden[f"{tag}_scores"]) and np.array_equal(r.ids, golden[f"{tag}_ids"])


@pytest.mark.gpu
def test_gpu_landmark_full_size_properties():
    """1M x 1536 with the default 1024 landmarks (the benchmark's index), where
    the oracle is too slow: a batch of queries equals the same queries one at
    a time bit for bit, device search equals host search, rows are sorted by
    (score desc, id asc) with ids in range, and a stored vector queried by
    itself ranks first."""
    import torch

    from vectorcache_amd import synthetic
    from vectorcache_amd.landmark import LandmarkIndex

    dev = torch.device("cuda:0")
    n, dim, k = 1_000_000, 1536, 10
    x = synthetic.anisotropic_torch(n, dim, 2001, dev)
    q = synthetic.anisotropic_torch(8, dim, 2003, dev)
    q[0] = x[123457]
    ix = LandmarkIndex(dim, 4)
    ix.add(x)
    ix.prepare()
    del x
    qh = q.cpu().numpy()
    batch = ix.search(qh, k)
    for j in range(qh.shape[0]):
        one = ix.search(qh[j: j + 1], k)
        assert bits_equal(one.scores[0], batch.scores[j]) and np.array_equal(one.ids[0], batch.ids[j])
    s, i = ix.search_device(q, k)
    torch.cuda.synchronize()
    assert bits_equal(s.cpu().numpy(), batch.scores) and np.array_equal(i.cpu().numpy().view(np.uint64), batch.ids)
    assert (batch.ids < n).all()
    for j in range(qh.shape[0]):
        sc, ids = batch.scores[j], batch.ids[j]
        assert all(sc[t] > sc[t + 1] or (sc[t] == sc[t + 1] and ids[t] < ids[t + 1]) for t in range(k - 1))
    assert batch.ids[0, 0] == 123457
