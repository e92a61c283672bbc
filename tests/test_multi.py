"""Many collections in one persistent scan launch (vc_multi_*) on the GPU:
uneven sizes, dims 128-1536, bits 2-4, batches of 1-1000 queries, k 1-100
(one k above a collection's size); every collection's rows equal its own
search_device bit for bit, eagerly under every plan, from graph replays with
new queries and eager work between them, and after the autotuner picks a
plan; the tuning table round-trips through its file and a single-index plan
measured by the tuner leaves search results unchanged."""
import numpy as np
import pytest

from util import bits_equal

pytestmark = pytest.mark.gpu

# (n, dim, bits, nq, k, calibrate)
SPECS = [
    (5_000, 128, 4, 1, 10, False),
    (30_000, 128, 2, 37, 100, False),
    (100_000, 768, 4, 300, 10, True),
    (64_000, 768, 3, 8, 1, False),
    (20_000, 1536, 4, 65, 100, True),
    (150_000, 1536, 2, 3, 10, True),
    (7, 128, 4, 5, 10, False),
    (45_000, 384, 4, 1000, 64, False),
]


@pytest.fixture(scope="module")
def collections(vc):
    import torch

    from vectorcache_amd import synthetic

    dev = torch.device("cuda:0")
    out = []
    for j, (n, dim, bits, nq, k, cal) in enumerate(SPECS):
        x = synthetic.zipf_mixture_torch(n, dim, 64, 0.0, 0.3, 3000 + j, 3100 + j, dev)
        ix = vc.TurboQuantIndex(dim, bits)
        if cal:
            ix.calibrate(synthetic.zipf_mixture_torch(2000, dim, 64, 0.0, 0.3, 3000 + j, 3200 + j, dev))
        ix.add(x)
        ix.prepare()
        qs = [synthetic.zipf_mixture_torch(nq, dim, 64, 0.0, 0.3, 3000 + j, 3300 + 10 * j + r, dev) for r in range(3)]
        out.append((ix, qs, k))
    torch.cuda.synchronize()
    return out


def _standalone(collections, r):
    res = []
    for ix, qs, k in collections:
        s, i = ix.search_device(qs[r], k)
        res.append((s.cpu().numpy(), i.cpu().numpy()))
    return res


def _check(ms, want, tag):
    import torch

    torch.cuda.synchronize()
    for j, (ws, wi) in enumerate(want):
        s, i = ms.results(j)
        s, i = s.cpu().numpy(), i.cpu().numpy()
        assert s.shape == ws.shape, (tag, j)
        assert bits_equal(s, ws) and np.array_equal(i, wi), (tag, j, SPECS[j])


def test_eager_equals_standalone_under_every_plan(collections):
    from vectorcache_amd.multi import MultiSearch

    want = _standalone(collections, 0)
    ms = MultiSearch([c[0] for c in collections], [s[3] for s in SPECS], [s[4] for s in SPECS])
    try:
        for j, (ix, qs, k) in enumerate(collections):
            ms.set_queries(j, qs[0])
        ms.search()
        _check(ms, want, "default")
        print(f"[multi] default plan {ms.plan()}", flush=True)
        for tile, vpt, scale, wg in ((8, 1, 0.5, 2), (4, 2, 2.0, 1), (2, 1, 1.0, 2), (1, 2, 4.0, 1), (8, 2, 0.25, 1),
                                     (8, 2, 1.0, 2)):
            ms.set_plan(tile, vpt, scale, wg)
            ms.search()
            _check(ms, want, (tile, vpt, scale, wg))
    finally:
        ms.close()


def test_graph_replays_equal_standalone(collections):
    import torch

    from vectorcache_amd.multi import MultiSearch

    wants = [_standalone(collections, r) for r in range(3)]
    ms = MultiSearch([c[0] for c in collections], [s[3] for s in SPECS], [s[4] for s in SPECS])
    side = torch.cuda.Stream()
    try:
        for j, (ix, qs, k) in enumerate(collections):
            ms.set_queries(j, qs[0])
        ms.search()  # eager run first
        torch.cuda.synchronize()
        ms.capture(side.cuda_stream)
        a = torch.randn(512, 512, device="cuda")
        for rep in range(6):
            r = rep % 3
            with torch.cuda.stream(side):
                for j, (ix, qs, k) in enumerate(collections):
                    ms.set_queries(j, qs[r], stream=side.cuda_stream)
                ms.replay(stream=side.cuda_stream)
            side.synchronize()
            _ = (a @ a).sum().item()  # eager work between replays
            _check(ms, wants[r], ("replay", rep))
    finally:
        ms.close()


def test_tuned_plan_and_tuning_file(collections, tmp_path):
    from vectorcache_amd import multi

    want = _standalone(collections, 1)
    ms = multi.MultiSearch([c[0] for c in collections], [s[3] for s in SPECS], [s[4] for s in SPECS])
    try:
        for j, (ix, qs, k) in enumerate(collections):
            ms.set_queries(j, qs[1])
        best = ms.tune(reps=2)
        print(f"[multi] tuned plan {best}", flush=True)
        assert best.measured_ms > 0
        ms.search()
        _check(ms, want, "tuned")
    finally:
        ms.close()
    path = tmp_path / "plans.txt"
    multi.save_tuning(path)
    n = multi.tuning_entries()
    assert n >= 1 and "multi sig=" in path.read_text()
    multi.clear_tuning()
    assert multi.tuning_entries() == 0
    assert multi.load_tuning(path) == n
    ms2 = multi.MultiSearch([c[0] for c in collections], [s[3] for s in SPECS], [s[4] for s in SPECS])
    try:
        p2 = ms2.plan()
        assert (p2.tile, p2.vec_per_thread, p2.item_scale, p2.workgroups_per_cu) == (
            best.tile, best.vec_per_thread, best.item_scale, best.workgroups_per_cu)
    finally:
        ms2.close()
    bad = tmp_path / "bad.txt"
    bad.write_text("scan g16=3 nlog=1 nqlog=1 kcls=1 tile=5 vpt=2 rscale=1 ms=1\n")
    with pytest.raises(Exception):
        multi.load_tuning(bad)
    multi.clear_tuning()


def test_scan_tuner_keeps_results(vc):
    import torch

    from vectorcache_amd import multi, synthetic

    dev = torch.device("cuda:0")
    n, dim, bits, nq, k = 200_000, 768, 4, 64, 10
    x = synthetic.zipf_mixture_torch(n, dim, 64, 0.0, 0.3, 3500, 3501, dev)
    q = synthetic.zipf_mixture_torch(nq, dim, 64, 0.0, 0.3, 3500, 3502, dev)
    ix = vc.TurboQuantIndex(dim, bits)
    ix.add(x)
    ix.prepare()
    multi.clear_tuning()
    s0, i0 = ix.search_device(q, k)
    choice = multi.tune_scan(dim, bits, n, nq, k, reps=2)
    print(f"[multi] scan tuner {choice}", flush=True)
    assert choice["ms"] <= choice["rule_ms"] and multi.tuning_entries() == 1
    s1, i1 = ix.search_device(q, k)
    assert bits_equal(s0.cpu().numpy(), s1.cpu().numpy()) and torch.equal(i0, i1)
    multi.clear_tuning()


def test_indexes_in_a_set_refuse_changes(vc):
    from vectorcache_amd._capi import IndexStateError
    from vectorcache_amd.multi import MultiSearch

    rng = np.random.default_rng(3600)
    x = rng.standard_normal((3000, 128)).astype(np.float32)
    ix = vc.TurboQuantIndex(128, 4)
    ix.add(x[:2000])
    ix.prepare()
    ms = MultiSearch([ix], [4], [10])
    with pytest.raises(IndexStateError):
        ix.add(x[2000:])
    with pytest.raises(IndexStateError):
        ix.close()
    with pytest.raises(ValueError):
        MultiSearch([ix, ix], [1, 1], [5, 5])  # one index twice
    ms.close()
    ix.add(x[2000:])
    ix.prepare()
    assert ix.size() == 3000
