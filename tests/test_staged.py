"""The landmark index's staged sign-plane search: the branch the reference
takes on an AVX-512 VNNI host once 32768 vectors are prepared
(src/search/search.cpp:1379-1510; sign planes src/pack/pack.cpp:428-541;
sign LUTs search.cpp:862-914; permute-dot rescore search_vnni.cpp:
1194-1360).

The oracle's restatement is pinned against golden vectors written from the
reference build running its VNNI kernels (tests/golden/landmark_staged.npz,
tests/golden/make_golden_staged.py) and, on a VNNI host with the reference
build present, against that build directly, with planted fusion mistakes that
must fail. Marked gpu: the HIP path through the C ABI against the golden
vectors and the oracle.

Bar: bit-exact scores and ids. The reference's buffered collector (sample
seeds, nth_element compaction, strict admission above the last kept score)
returns exactly the E best sign scores by (score desc, id asc), ties included,
because one query's vectors reach it in ascending id order:
test_buffered_collector_returns_exact_cut replays the collector on tied
scores. Its nth_element leaves the shortlist's order unspecified, which could
matter only for two candidates tying exactly on the final score at the k-th
place.

Scan widths that are not a multiple of 32 have no sign kernel in the reference
(search_vnni.cpp:1466-1469 falls through to the 192-wide 4-bit kernel, which
reads past the 32-wide sign blocks): there the planes index keeps the exact
scan, the AVX2 host's result.
"""
import numpy as np
import pytest

from cases import STAGED_CASES, UNSTAGED_PLANES_CASE, landmark_inputs
from conftest import load_golden
from util import scores_equal


def _host_has_vnni():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return all(f in flags for f in ("avx512_vnni", "avx512vbmi", "avx512bw"))


@pytest.fixture(scope="module")
def golden():
    return load_golden("landmark_staged.npz")


def _run(o, case, **kw):
    tag, n, nq, dim, m, seed, k = case
    x, q = landmark_inputs(n, nq, dim)
    return o.lm_search(x, q, k, scan_bits=4, n_landmarks=m, seed=seed, planes=True, **kw)


# ---- oracle (CPU) ------------------------------------------------------------

@pytest.mark.parametrize("case", STAGED_CASES, ids=[c[0] for c in STAGED_CASES])
def test_oracle_staged_matches_golden(oracle, golden, case):
    s, i, _ = _run(oracle, case)
    assert np.array_equal(i, golden[f"{case[0]}_ids"])
    assert scores_equal(s, golden[f"{case[0]}_scores"])


def _collector_cut(scores, E, seed=-np.inf):
    """The buffered collector of search_vnni.cpp:368-500 and topk_heap.hpp:
    66-117 on one query's scores in id order: append when score > thr; at 2E
    entries keep the best E by (score desc, id asc) and set thr to the least
    kept score; the final shortlist is the best E of what is left
    (search.cpp:569-592); a seeded run short of E is redone unseeded
    (search.cpp:1411-1440)."""
    def run(thr):
        buf = []
        for i, sc in enumerate(scores):
            if not sc > thr:
                continue
            buf.append((-sc, i))
            if len(buf) == 2 * E:
                buf = sorted(buf)[:E]
                thr = -buf[-1][0]
        return sorted(buf)[:E]

    out = run(seed)
    if len(out) < E:
        out = run(-np.inf)
    return [i for _, i in out]


@pytest.mark.parametrize("levels", [3, 9, 40, 100000])
def test_buffered_collector_returns_exact_cut(levels):
    """Heavily tied scores (few levels) and seeds below, at and above the
    E-th score: the collector's shortlist is the exact top E."""
    rng = np.random.default_rng(levels)
    for trial in range(20):
        n = int(rng.integers(50, 400))
        E = int(rng.integers(1, n // 2 + 1))
        scores = rng.integers(0, levels, n).astype(np.float64) * 0.25 - 3.0
        exact = [i for _, i in sorted((-sc, i) for i, sc in enumerate(scores))[:E]]
        kth = scores[exact[-1]]
        for seed in (-np.inf, kth - 0.25, kth, kth + 0.25):
            assert _collector_cut(scores, E, seed) == exact, (levels, trial, seed)


def test_oracle_unstaged_width_keeps_exact_scan(oracle):
    """n_landmarks = 200 (scan_dim % 32 != 0): planes give the exact scan."""
    tag, n, nq, dim, m, seed, k = UNSTAGED_PLANES_CASE
    x, q = landmark_inputs(n, nq, dim)
    s_p, i_p, _ = oracle.lm_search(x, q, k, scan_bits=4, n_landmarks=m, seed=seed, planes=True)
    s_a, i_a, _ = oracle.lm_search(x, q, k, scan_bits=4, n_landmarks=m, seed=seed)
    assert np.array_equal(i_p, i_a) and np.array_equal(s_p.view(np.uint32), s_a.view(np.uint32))


def test_oracle_staged_shortlist_lengths(oracle):
    """planes_exact_len(k, planes_shortlist_len(k, 4)) (pack.cpp:451-463)."""
    assert [oracle.planes_exact_len(k) for k in (1, 10, 42, 43, 63, 64, 100, 341)] == \
        [256, 256, 256, 258, 378, 384, 600, 2046]


def test_oracle_staged_fusion_is_pinned(oracle, ref):
    """Against the reference build's VNNI kernels: the sign scan (fused
    A * scale + bias) and the whole staged search equal the oracle; the
    unfused sign score and the other fusion of the rescore flush do not."""
    if not _host_has_vnni():
        pytest.skip("the reference's VNNI kernels need an AVX-512 VNNI/VBMI host")
    rng = np.random.default_rng(5)
    n, S = 5000, 64
    packed = rng.integers(0, 256, (n, 4 * S // 8), dtype=np.uint8)
    rows = (rng.standard_normal((3, S)) * 0.1).astype(np.float32)
    sr, ir = ref.sign_topk(rows, packed, 40)
    so, io = oracle.sign_topk(rows, packed, 40)
    assert np.array_equal(ir, io) and np.array_equal(sr.view(np.uint32), so.view(np.uint32))
    try:
        oracle.set_staged_variant(1)
        su, _ = oracle.sign_topk(rows, packed, 40)
        assert not np.array_equal(sr.view(np.uint32), su.view(np.uint32)), "unfused sign scores must differ"
        oracle.set_staged_variant(2)
        case = ("extra", 34000, 5, 128, 96, 21, 20)
        sp, ip, _ = _run(oracle, case)
        oracle.set_staged_variant(0)
        s0, i0, _ = _run(oracle, case)
        s1, i1, _ = _run(ref, case)
        assert np.array_equal(i0, i1) and np.array_equal(s0.view(np.uint32), s1.view(np.uint32))
        assert not np.array_equal(sp.view(np.uint32), s1.view(np.uint32)), "the other flush fusion must differ"
    finally:
        oracle.set_staged_variant(0)


# ---- HIP path (GPU) ------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("case", STAGED_CASES, ids=[c[0] for c in STAGED_CASES])
def test_gpu_staged_matches_golden(golden, case):
    from vectorcache_amd.landmark import LandmarkIndex

    tag, n, nq, dim, m, seed, k = case
    x, q = landmark_inputs(n, nq, dim)
    ix = LandmarkIndex(dim, 4, m, seed, planes=True)
    ix.add(x)
    ix.prepare()
    assert ix.planes_active()
    r = ix.search(q, k)
    assert np.array_equal(r.ids, golden[f"{tag}_ids"]), tag
    assert scores_equal(r.scores, golden[f"{tag}_scores"]), tag


@pytest.mark.gpu
def test_gpu_staged_matches_oracle_and_device_path(oracle):
    """Another shape against the oracle; device search equals host search;
    switching planes off gives the AVX2 branch's results."""
    import torch

    from vectorcache_amd.landmark import LandmarkIndex

    case = ("gpu", 38000, 9, 192, 128, 4, 32)
    tag, n, nq, dim, m, seed, k = case
    x, q = landmark_inputs(n, nq, dim)
    s_o, i_o, _ = _run(oracle, case)
    ix = LandmarkIndex(dim, 4, m, seed, planes=True)
    ix.add(x)
    ix.prepare()
    r = ix.search(q, k)
    assert np.array_equal(r.ids, i_o) and scores_equal(r.scores, s_o)
    sd, idd = ix.search_device(torch.from_numpy(q).cuda(), k)
    torch.cuda.synchronize()
    assert np.array_equal(idd.cpu().numpy().astype(np.uint64), r.ids)
    assert scores_equal(sd.cpu().numpy(), r.scores)
    ix.set_planes(False)
    assert not ix.planes_active()
    s_a, i_a, _ = oracle.lm_search(x, q, k, scan_bits=4, n_landmarks=m, seed=seed)
    r2 = ix.search(q, k)
    assert np.array_equal(r2.ids, i_a) and scores_equal(r2.scores, s_a)


@pytest.mark.gpu
def test_gpu_planes_below_threshold_keep_the_avx2_scan(oracle):
    """Below 32768 prepared vectors no sign planes are kept (pack.cpp:
    428-449): the planes index searches as the AVX2 branch does."""
    from vectorcache_amd.landmark import LandmarkIndex

    x, q = landmark_inputs(3000, 4, 128)
    ix = LandmarkIndex(128, 4, 64, 3, planes=True)
    ix.add(x)
    ix.prepare()
    assert not ix.planes_active()
    s_a, i_a, _ = oracle.lm_search(x, q, 10, scan_bits=4, n_landmarks=64, seed=3)
    r = ix.search(q, 10)
    assert np.array_equal(r.ids, i_a) and scores_equal(r.scores, s_a)


@pytest.mark.gpu
def test_gpu_planes_unstaged_width_keeps_the_exact_scan(oracle):
    """scan_dim % 32 != 0: no sign planes are kept and the search equals the
    oracle's exact scan."""
    from vectorcache_amd.landmark import LandmarkIndex

    tag, n, nq, dim, m, seed, k = UNSTAGED_PLANES_CASE
    x, q = landmark_inputs(n, nq, dim)
    ix = LandmarkIndex(dim, 4, m, seed, planes=True)
    ix.add(x)
    ix.prepare()
    assert not ix.planes_active()
    s_a, i_a, _ = oracle.lm_search(x, q, k, scan_bits=4, n_landmarks=m, seed=seed)
    r = ix.search(q, k)
    assert np.array_equal(r.ids, i_a) and scores_equal(r.scores, s_a)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 341])
def test_gpu_staged_k_limits_match_oracle(oracle, k):
    """k = 1 (E = 256) and k = 341, the largest k whose shortlist fits the
    GPU top-k (E = 2046): against the oracle."""
    from vectorcache_amd.landmark import LandmarkIndex

    case = ("klim", 34000, 3, 128, 64, 2, k)
    tag, n, nq, dim, m, seed, kk = case
    x, q = landmark_inputs(n, nq, dim)
    s_o, i_o, _ = _run(oracle, case)
    ix = LandmarkIndex(dim, 4, m, seed, planes=True)
    ix.add(x)
    ix.prepare()
    r = ix.search(q, kk)
    assert np.array_equal(r.ids, i_o) and scores_equal(r.scores, s_o)


@pytest.mark.gpu
def test_gpu_staged_rejects_a_shortlist_past_the_gpu_top_k():
    """k = 342 needs a 2052-row shortlist: a clear error, and the index keeps
    working."""
    from vectorcache_amd.landmark import LandmarkIndex

    x, q = landmark_inputs(34000, 2, 128)
    ix = LandmarkIndex(128, 4, 64, 2, planes=True)
    ix.add(x)
    ix.prepare()
    with pytest.raises(Exception, match="shortlist"):
        ix.search(q, 342)
    assert ix.search(q, 10).ids.shape == (2, 10)


@pytest.mark.gpu
def test_gpu_staged_batches_and_empty_queries(oracle):
    """Adds in three batches with a prepare after each (the sign planes are
    rebuilt every time; the first prepare is below 32768 vectors, so no planes
    yet), then no queries at all."""
    from vectorcache_amd.landmark import LandmarkIndex

    n, nq, dim, m, seed, k = 36000, 4, 128, 64, 5, 20
    x, q = landmark_inputs(n, nq, dim)
    batches = [20000, 10000, 6000]
    ix = LandmarkIndex(dim, 4, m, seed, planes=True)
    start = 0
    for b in batches:
        ix.add(x[start:start + b])
        ix.prepare()
        start += b
        assert ix.planes_active() == (start >= 32768)
    s_o, i_o, _ = oracle.lm_search(x, q, k, scan_bits=4, n_landmarks=m, seed=seed, batches=batches,
                                   prepare_between=True, planes=True)
    r = ix.search(q, k)
    assert np.array_equal(r.ids, i_o) and scores_equal(r.scores, s_o)
    r0 = ix.search(q[:0], k)
    assert r0.ids.shape[0] == 0
