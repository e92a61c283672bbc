"""GPU tests of the bucketed index (vc_bucketed_*) and its learned probe model.

Parity: every bucket's codes and alpha equal the oracle's encode of the
oracle's residual rows; a search's scores equal, bit for bit, the oracle's
chain of score_prepared_into calls over the same probed buckets with the same
offsets (ids equal wherever a score is not tied with the k-th one: the heap
chain keeps the first of equal scores, the GPU the lower id). The bucket
assignment and the routing choose buckets only and are parity-unpinned; the
tests take the GPU's bucket lists as the oracle's input.
The probe model's gradients are checked against the CPU fp64 oracle
(oracle/probe_oracle.py) and its Adam step against the oracle's update.
"""
import numpy as np
import pytest

from util import bits_equal, scores_equal

pytestmark = pytest.mark.gpu


def _bucketed(x, dim, bits, residual=True, sample=None, **kw):
    from vectorcache_amd.bucketed import BucketedTurboQuantIndex, BucketParams

    ix = BucketedTurboQuantIndex(dim, bits, BucketParams(residual=residual, **kw))
    if sample is not None:
        ix.calibrate(sample)
    ix.add(x)
    ix.prepare()
    return ix


def _pos_state(st):
    sizes = st["bucket_sizes"].astype(np.int64)
    pos_bucket = np.repeat(np.arange(sizes.size), sizes).astype(np.uint32)
    return sizes, pos_bucket


def _same_topk(got_s, got_i, want_s, want_i):
    """Scores bit-equal; ids equal except where the score ties the row's last one."""
    assert scores_equal(got_s, want_s)
    for r in range(got_s.shape[0]):
        last = want_s[r, -1]
        keep = want_s[r] != last
        assert np.array_equal(got_i[r][keep], want_i[r][keep]), r


CASES = [
    # (dim, bits, n, energy_soft, max_bucket_size, tq)
    (64, 4, 6000, 100, 200, False),
    (128, 2, 5000, 120, 240, True),
    (96, 3, 4000, 80, 160, False),
]


@pytest.mark.parametrize("dim,bits,n,es,mb,tq", CASES)
def test_bucket_codes_match_oracle(vc, oracle, dim, bits, n, es, mb, tq):
    from vectorcache_amd import synthetic

    x = synthetic.zipf_mixture(n, dim, 40, 1.1, 0.3, 11, 12)
    sample = synthetic.zipf_mixture(600, dim, 40, 1.1, 0.3, 11, 13) if tq else None
    ix = _bucketed(x, dim, bits, sample=sample, energy_soft=es, max_bucket_size=mb)
    st = ix.export()
    sizes, pos_bucket = _pos_state(st)
    assert sizes.sum() == n and sizes.max() <= mb and ix.n_buckets() == sizes.size
    assert np.array_equal(np.sort(st["pos_ids"]), np.arange(n, dtype=np.uint64))
    res = oracle.bucket_residuals(x, st["pos_ids"], pos_bucket, st["centroids"])
    sh, sc = (st["shift"], st["scale"]) if tq else (None, None)
    packed, alpha = oracle.encode(res, bits, sh, sc)
    assert np.array_equal(oracle.codes_flat(packed, bits, dim), st["flat"])
    assert bits_equal(alpha, st["alpha"])
    # a row's bucket is its nearest centroid up to bf16 routing and the 2-means splits
    xn = x / np.linalg.norm(x, axis=1, keepdims=True)
    rows = xn[st["pos_ids"].astype(np.int64)]
    own = (rows * st["centroids"][pos_bucket]).sum(1)
    best = (rows @ st["centroids"].T).max(1)
    assert np.mean(own >= best - 0.05) > 0.95


# debug flags: bit 7 forces the bucket-major tile scan, bit 8 the query-major one
SCAN_PATHS = {"query": 256, "tile": 128}


@pytest.fixture
def scan_path(vc, request):
    vc.set_debug(SCAN_PATHS[request.param])
    yield request.param
    vc.set_debug(0)


@pytest.mark.parametrize("scan_path", list(SCAN_PATHS), indirect=True)
@pytest.mark.parametrize("dim,bits,n,es,mb,tq", CASES)
def test_bucketed_search_matches_oracle_chain(vc, oracle, dim, bits, n, es, mb, tq, scan_path):
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.zipf_mixture(n, dim, 40, 1.1, 0.3, 11, 12)
    q = synthetic.zipf_mixture(24, dim, 40, 1.1, 0.3, 11, 14)
    sample = synthetic.zipf_mixture(600, dim, 40, 1.1, 0.3, 11, 13) if tq else None
    ix = _bucketed(x, dim, bits, sample=sample, energy_soft=es, max_bucket_size=mb)
    st = ix.export()
    L = ix.n_buckets()
    qd = torch.from_numpy(q).cuda()
    sh, sc = (st["shift"], st["scale"]) if tq else (None, None)
    for n_probe in (1, 5, L):
        P = min(n_probe, L)
        pb, ps = ix.route_device(qd, P)
        pb = pb.cpu().numpy()
        for k in (10, 64):
            s, i, cnt = ix.search_device(qd, k, n_probe=n_probe, counts=torch.empty((24, 1), dtype=torch.int64,
                                                                                     device="cuda"))
            s, i, cnt = s.cpu().numpy(), i.cpu().numpy().astype(np.uint64), cnt.cpu().numpy().ravel()
            ws, wi = oracle.bucketed_search(q, k, bits, st["bucket_sizes"], st["flat"], st["alpha"], st["pos_ids"],
                                            st["centroids"], pb, np.full(24, P, np.int32), sh, sc)
            found = np.array([min(k, int(st["bucket_sizes"][pb[r]].sum())) for r in range(24)])
            assert np.array_equal(cnt, found), (n_probe, k)
            _same_topk(s, i, ws, wi)


def _tile_vs_query(vc, ix, qd, k, **kw):
    import torch

    nq = qd.shape[0]
    out = {}
    for name, flag in SCAN_PATHS.items():
        vc.set_debug(flag)
        try:
            s, i, c = ix.search_device(qd, k, counts=torch.empty((nq, 1), dtype=torch.int64, device="cuda"), **kw)
        finally:
            vc.set_debug(0)
        out[name] = (s.cpu().numpy(), i.cpu().numpy(), c.cpu().numpy().ravel())
    (qs, qi, qc), (ts, ti, tc) = out["query"], out["tile"]
    # both keep (score desc, position asc): ids equal even on ties
    assert scores_equal(ts, qs)
    assert np.array_equal(ti, qi)
    assert np.array_equal(tc, qc)
    return out["tile"]


@pytest.mark.parametrize("bits,tq", [(4, True), (2, False)])
def test_tile_scan_equals_query_scan(vc, oracle, bits, tq):
    """The bucket-major tile scan (eight queries per bucket read) returns the
    query-major scan's lists bit for bit: many queries per bucket, fixed and
    per-query probe counts, k from 1 to the tile limit of 64, duplicated rows
    for heavy score ties; 48 queries against the oracle's chain."""
    import torch

    from vectorcache_amd import synthetic

    dim, n = 128, 12000
    x = synthetic.zipf_mixture(n, dim, 30, 1.1, 0.3, 41, 42)
    x[6000:9000] = x[:3000]  # equal rows: equal scores, broken by position
    q = synthetic.zipf_mixture(700, dim, 30, 1.1, 0.3, 41, 43)
    sample = synthetic.zipf_mixture(600, dim, 30, 1.1, 0.3, 41, 44) if tq else None
    ix = _bucketed(x, dim, bits, sample=sample, energy_soft=100, max_bucket_size=200)
    L = ix.n_buckets()
    qd = torch.from_numpy(q).cuda()
    rng = np.random.default_rng(45)
    for n_probe in (3, 17, L):
        for k in (1, 10, 64):
            _tile_vs_query(vc, ix, qd, k, n_probe=n_probe)
    counts = torch.from_numpy(rng.integers(0, 40, 700).astype(np.int32)).cuda()
    ts, ti, _ = _tile_vs_query(vc, ix, qd, 10, n_probe=40, probe_counts=counts)
    # the oracle's chain over each query's first counts[q] buckets
    st = ix.export()
    pb = ix.route_device(qd[:48], 40)[0].cpu().numpy()
    sh, sc = (st["shift"], st["scale"]) if tq else (None, None)
    cnt = np.clip(counts[:48].cpu().numpy(), 1, 40).astype(np.int32)  # counts are clamped to [1, n_probe]
    ws, wi = oracle.bucketed_search(q[:48], 10, bits, st["bucket_sizes"], st["flat"], st["alpha"], st["pos_ids"],
                                    st["centroids"], pb, cnt, sh, sc)
    # rows with fewer than k vectors opened hold (0, 0) past their count in both
    _same_topk(ts[:48], ti[:48].astype(np.uint64), ws, wi)


def test_tile_scan_buckets_above_one_pass(vc, oracle):
    """Buckets of more than 1024 vectors take several passes of the tile scan
    over the bucket (the lists carried across the passes): the same lists as
    the query-major scan, and as the oracle's chain for 16 queries."""
    import torch

    from vectorcache_amd import synthetic

    dim, bits, n = 64, 4, 9000
    x = synthetic.zipf_mixture(n, dim, 6, 1.1, 0.3, 51, 52)
    q = synthetic.zipf_mixture(300, dim, 6, 1.1, 0.3, 51, 53)
    ix = _bucketed(x, dim, bits, energy_soft=1500, max_bucket_size=2600)
    st = ix.export()
    assert st["bucket_sizes"].max() > 2048  # three passes of 1024
    L = ix.n_buckets()
    qd = torch.from_numpy(q).cuda()
    for k in (1, 10, 64):
        ts, ti, _ = _tile_vs_query(vc, ix, qd, k, n_probe=min(3, L))
    pb = ix.route_device(qd[:16], min(3, L))[0].cpu().numpy()
    ws, wi = oracle.bucketed_search(q[:16], 64, bits, st["bucket_sizes"], st["flat"], st["alpha"], st["pos_ids"],
                                    st["centroids"], pb, np.full(16, min(3, L), np.int32))
    _same_topk(ts[:16], ti[:16].astype(np.uint64), ws, wi)


def test_bucketed_direct_codes_all_buckets_equal_flat_search(vc, oracle):
    """residual = 0: codes of the normalized rows; opening every bucket gives
    the flat index's search of the same rows, ties included."""
    import torch

    from vectorcache_amd import synthetic

    dim, bits, n = 128, 4, 5000
    x = synthetic.zipf_mixture(n, dim, 40, 1.1, 0.3, 21, 22)
    q = synthetic.zipf_mixture(16, dim, 40, 1.1, 0.3, 21, 23)
    ix = _bucketed(x, dim, bits, residual=False, energy_soft=100, max_bucket_size=200)
    xn = np.stack([r * (np.float32(1) / oracle.row_norm(r)) for r in x]).astype(np.float32)
    flat = vc.TurboQuantIndex(dim, bits)
    flat.add(xn)
    flat.prepare()
    for k in (10, 100):
        want = flat.search(q, k)
        s, i, _ = ix.search_device(torch.from_numpy(q).cuda(), k, n_probe=ix.n_buckets())
        assert scores_equal(s.cpu().numpy(), want.scores)
        assert np.array_equal(i.cpu().numpy().astype(np.uint64), want.ids)


def test_bucketed_edge_cases(vc):
    import torch

    from vectorcache_amd import synthetic
    from vectorcache_amd.bucketed import BucketedTurboQuantIndex, BucketParams

    with pytest.raises(ValueError):
        BucketedTurboQuantIndex(100, 4)
    with pytest.raises(ValueError):
        BucketedTurboQuantIndex(64, 5)
    with pytest.raises(ValueError):
        BucketedTurboQuantIndex(64, 4, BucketParams(energy_soft=900))
    ix = BucketedTurboQuantIndex(64, 4, BucketParams(energy_soft=50, max_bucket_size=100))
    from vectorcache_amd._capi import IndexStateError

    with pytest.raises(IndexStateError):
        ix.search(np.zeros((1, 64), np.float32), 5)
    x = synthetic.zipf_mixture(30, 64, 4, 1.1, 0.3, 31, 32)
    ix.add(x)
    ix.prepare()
    # fewer vectors than k: rows hold every opened vector, then (0, 0)
    r = ix.search(x[:3], 50)
    assert r.ids.shape == (3, 30)
    s, i, cnt = ix.search_device(torch.from_numpy(x[:3]).cuda(), 50, n_probe=1,
                                 counts=torch.empty((3, 1), dtype=torch.int64, device="cuda"))
    cnt = cnt.cpu().numpy().ravel()
    assert (cnt >= 1).all() and (cnt <= 30).all()
    # a row queried by itself ranks first
    assert np.array_equal(r.ids[:, 0], np.arange(3, dtype=np.uint64))
    # add after prepare rebuilds at the next prepare
    ix.add(x[:10])
    assert not ix.prepared
    ix.prepare()
    assert ix.size() == 40


def test_probe_features_and_counts(vc):
    """Features of the routing lists; per-query probe counts are clamped and
    honoured (a count of c opens exactly the c best buckets)."""
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.zipf_mixture(8000, 64, 30, 1.1, 0.3, 41, 42)
    q = synthetic.zipf_mixture(32, 64, 30, 1.1, 0.3, 41, 43)
    ix = _bucketed(x, 64, 4, energy_soft=100, max_bucket_size=200)
    qd = torch.from_numpy(q).cuda()
    L = ix.n_buckets()
    P = min(256, L)
    pb, ps = ix.route_device(qd, P)
    ps = ps.cpu().numpy()
    f = ix.probe_features(qd).cpu().numpy()
    sizes = ix.export()["bucket_sizes"]
    assert np.allclose(f[:, 0], ps[:, 0])
    assert np.allclose(f[:, 1:64], 16 * (ps[:, :1] - ps[:, 1:64]), atol=1e-5)
    assert np.allclose(f[:, 64:], np.log2(1 + sizes[pb.cpu().numpy()[:, :32]]) / 10, atol=1e-6)
    counts = torch.from_numpy(np.arange(32, dtype=np.int32) % 7).cuda()  # 0 clamps to 1
    s, i, c = ix.search_device(qd, 10, n_probe=6, probe_counts=counts)
    for r in range(32):
        want = min(max(r % 7, 1), 6)
        s1, i1, _ = ix.search_device(qd[r:r + 1], 10, n_probe=want)
        assert torch.equal(i[r], i1[0]) and torch.equal(s[r], s1[0]), r


def _probe_model(ix, seed=5):
    import probe_oracle as po

    ix.probe_init(seed)
    return ix.probe_weights()


def test_probe_gradients_match_cpu_oracle(vc):
    """One minibatch's gradients from the GPU kernels against the fp64 oracle
    with the same bf16 rounding points (relative error per tensor), and the
    plain fp64 gradients within the bf16 effect."""
    import probe_oracle as po
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.zipf_mixture(3000, 64, 20, 1.1, 0.3, 51, 52)
    ix = _bucketed(x, 64, 4, energy_soft=60, max_bucket_size=120)
    w = _probe_model(ix)
    r = np.random.default_rng(53)
    X = (r.standard_normal((200, 96)) * 0.7).astype(np.float32)
    t = r.uniform(0, 8, 200).astype(np.float32)
    order = r.permutation(200)[:64].astype(np.int32)
    g = ix.probe_grads(torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda(), torch.from_numpy(order).cuda())
    want = po.grads(w, X[order], t[order], bf16_points=True)
    plain = po.grads(w, X[order], t[order], bf16_points=False)
    for lo, hi in ((po.O_W1, po.O_B1), (po.O_B1, po.O_W2), (po.O_W2, po.O_B2), (po.O_B2, po.O_W3), (po.O_W3, po.O_B3),
                   (po.O_B3, po.O_B3 + 1)):
        err = np.linalg.norm(g[lo:hi] - want[lo:hi]) / np.linalg.norm(want[lo:hi])
        assert err < 2e-3, (lo, err)
        assert np.linalg.norm(g[lo:hi] - plain[lo:hi]) / np.linalg.norm(plain[lo:hi]) < 0.15, lo


def test_probe_adam_step_matches_oracle(vc):
    """One training step: the new weights are the oracle's Adam update of the
    old weights with the GPU's own gradients (fp32 vs fp64 arithmetic)."""
    import probe_oracle as po
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.zipf_mixture(3000, 64, 20, 1.1, 0.3, 51, 52)
    ix = _bucketed(x, 64, 4, energy_soft=60, max_bucket_size=120)
    w0 = _probe_model(ix, 7)
    r = np.random.default_rng(54)
    X = (r.standard_normal((64, 96)) * 0.7).astype(np.float32)
    t = r.uniform(0, 8, 64).astype(np.float32)
    order = np.arange(64, dtype=np.int32)
    Xd, td, od = torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda(), torch.from_numpy(order).cuda()
    g = ix.probe_grads(Xd, td, od)
    ix.probe_train(Xd, td, od, lr=1e-3)
    torch.cuda.synchronize()
    w1 = ix.probe_weights()
    want, _, _ = po.adam_step(w0, np.zeros_like(g), np.zeros_like(g), g.astype(np.float64), 1, 1e-3)
    live = slice(0, po.O_B3 + 1)
    assert np.allclose(w1[live], want[live], rtol=1e-5, atol=1e-7)


def test_probe_training_lowers_loss_and_predicts(vc):
    """The GPU training loop fits a smooth target of the features; predict()
    equals the fused probe counts' model output."""
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.zipf_mixture(3000, 64, 20, 1.1, 0.3, 51, 52)
    ix = _bucketed(x, 64, 4, energy_soft=60, max_bucket_size=120)
    ix.probe_init(3)
    r = np.random.default_rng(55)
    X = r.uniform(-1, 1, (4096, 96)).astype(np.float32)
    t = (4 + 2 * np.tanh(X[:, 0] + X[:, 5] - X[:, 70])).astype(np.float32)
    Xd, td = torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda()
    first = None
    for ep in range(6):
        order = torch.from_numpy(r.permutation(4096).astype(np.int32)).cuda()
        losses = ix.probe_train(Xd, td, order, lr=2e-3).cpu().numpy()
        first = losses[:8].mean() if first is None else first
    last = losses[-8:].mean()
    assert last < 0.1 * first, (first, last)
    y = ix.probe_predict(Xd).cpu().numpy()
    assert np.mean((y - t) ** 2) < 0.1


def test_probe_table_counts_match_explicit_counts(vc):
    """A learned search with a probe table opens, per query, the table's count
    for the bin of the model's prediction: its results equal a search with
    those counts given explicitly."""
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.zipf_mixture(8000, 64, 30, 1.1, 0.3, 61, 62)
    q = synthetic.zipf_mixture(48, 64, 30, 1.1, 0.3, 61, 63)
    ix = _bucketed(x, 64, 4, energy_soft=100, max_bucket_size=200)
    ix.probe_init(9)
    ix.set_probe_weights(ix.probe_weights())  # given weights count as a trained model
    qd = torch.from_numpy(q).cuda()
    y = ix.probe_predict(ix.probe_features(qd)).cpu().numpy()
    table = np.array([1, 2, 3, 5, 8, 13, 21, 34], np.int32)
    y0, step = float(np.percentile(y, 10)), float((y.max() - y.min()) / 6 + 1e-3)
    ix.probe_set_table(table, y0, step)
    ix.probe_enable(True)
    L = ix.n_buckets()
    want = np.minimum(table[np.clip(np.floor((y - y0) / step), 0, len(table) - 1).astype(int)], min(256, L))
    s, i, _ = ix.search_device(qd, 10)
    s2, i2, _ = ix.search_device(qd, 10, n_probe=256, probe_counts=torch.from_numpy(want.astype(np.int32)).cuda())
    assert torch.equal(i, i2) and torch.equal(s, s2)
    assert len(set(want.tolist())) > 2


def test_routing_top_p_is_the_best_buckets(vc, oracle):
    """Routing lists on an index with more buckets than the selection's
    candidate buffer (the sampled-threshold path): sorted, and their scores
    agree with fp64 scores of the same bf16-rounded inputs; no bucket left out
    scores above the last one kept (up to the fp32 accumulation's rounding)."""
    import probe_oracle as po
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.zipf_mixture(70000, 64, 500, 1.1, 0.3, 71, 72)
    q = synthetic.zipf_mixture(40, 64, 500, 1.1, 0.3, 71, 73)
    q[0] = 0.0  # a zero query: every score ties (the exact fallback)
    ix = _bucketed(x, 64, 4, energy_soft=25, max_bucket_size=50, kmeans_iters=2)
    L = ix.n_buckets()
    assert L > 2048
    cent = ix.export()["centroids"]
    qn = np.stack([r * (np.float32(1) / oracle.row_norm(r) if oracle.row_norm(r) > 1e-10 else np.float32(0))
                   for r in q]).astype(np.float32)
    ref = po.bf16(qn).astype(np.float64) @ po.bf16(cent).astype(np.float64).T
    for P in (1, 64, 256, 512):
        pb, ps = ix.route_device(torch.from_numpy(q).cuda(), P)
        pb, ps = pb.cpu().numpy(), ps.cpu().numpy()
        for r in range(40):
            assert np.all(np.diff(ps[r]) <= 0), (P, r)
            assert len(set(pb[r].tolist())) == P
            assert np.allclose(ps[r], ref[r, pb[r]], atol=2e-5), (P, r)
            rest = np.setdiff1d(np.arange(L), pb[r])
            assert ref[r, rest].max() <= ps[r, -1] + 2e-5, (P, r)
        # the zero query: all scores 0, the lowest bucket ids first
        assert np.array_equal(pb[0], np.arange(P)) and np.all(ps[0] == 0)
