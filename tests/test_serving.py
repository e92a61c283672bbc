"""The request server (vc_server_*) on the GPU: requests submitted from
several threads while add batches grow the index are batched together
(continuous batching, admission by a small device-memory pool, a capacity
growth in the middle), and every request's result is bit-identical to the
same request searched alone over the vectors its batch saw; a few of them
also against the oracle's search_flat over those vectors' codes. Then the
admission refusals and the index's refusals while served."""
import threading
import time

import numpy as np
import pytest

from util import bits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(vc):
    import torch

    from vectorcache_amd import synthetic
    from vectorcache_amd.serving import Server, ServerParams, batch_bytes

    dev = torch.device("cuda:0")
    n0, n_add, dim = 200_000, 40_000, 384
    x = synthetic.anisotropic_torch(n0 + n_add, dim, 5101, dev)
    cal = synthetic.anisotropic_torch(5_000, dim, 5102, dev)
    ix = vc.TurboQuantIndex(dim, 4)
    ix.calibrate(cal)
    ix.add(x[:n0])
    ix.prepare()
    xa = x[n0:].cpu().numpy()
    qpool = synthetic.anisotropic_torch(4096, dim, 5103, dev).cpu().numpy()
    reserve = n0 + 20_000  # the adds outgrow it: one reallocation while serving
    worst = max(batch_bytes(ix, n, 256, 100) for n in (n0, n0 + n_add))
    print(f"[serving] one 256-query k=100 batch holds {worst} bytes", flush=True)
    params = ServerParams(slots=4, max_batch_queries=256, max_k=100, hbm_budget_bytes=int(worst * 2.5),
                          reserve_vectors=reserve, add_chunk=3000)
    sv = Server(ix, params)

    reqs, adds = [], []
    lock = threading.Lock()
    errors = []

    def submitter(seed, n_req):
        rng = np.random.default_rng(seed)
        mine = []
        try:
            for _ in range(n_req):
                nq = int(rng.integers(1, 65))
                q0 = int(rng.integers(0, qpool.shape[0] - nq))
                k = int(rng.choice([10, 100]))
                q = qpool[q0:q0 + nq]
                mine.append((sv.submit_search(q, k), q0, nq, k))
                time.sleep(float(rng.exponential(0.0005)))
            got = [(q0, nq, k, sv.wait(t)) for t, q0, nq, k in mine]
            with lock:
                reqs.extend(got)
        except Exception as e:  # surfaced by the test
            errors.append(e)

    def adder():
        try:
            ts = []
            for s in range(0, n_add, 5000):
                ts.append(sv.submit_add(xa[s:s + 5000]))
                time.sleep(0.004)
            adds.extend(sv.wait(t) for t in ts)
        except Exception as e:
            errors.append(e)

    threads = [threading.Thread(target=submitter, args=(5200 + i, 120)) for i in range(3)]
    threads.append(threading.Thread(target=adder))
    t0 = time.monotonic()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    print(f"[serving] {len(reqs)} requests, {len(adds)} adds in {time.monotonic() - t0:.2f} s", flush=True)
    stats = sv.stats()
    sv.close()
    assert not errors, errors
    return ix, qpool, xa, n0, reqs, adds, stats


def test_requests_equal_alone(served):
    import torch

    ix, qpool, xa, n0, reqs, adds, stats = served
    print(f"[serving] stats {stats}", flush=True)
    assert len(reqs) == 360 and len(adds) == 8
    assert ix.size() == n0 + xa.shape[0]
    assert stats["batches"] < stats["requests"]  # requests shared batches
    assert stats["vectors_added"] == xa.shape[0]
    assert stats["pool_peak_bytes"] <= stats["pool_bytes"]
    snaps = sorted({r[3].n_snapshot for r in reqs})
    assert snaps[0] >= n0 and snaps[-1] <= n0 + xa.shape[0]
    print(f"[serving] {len(snaps)} distinct snapshots, batch sizes up to "
          f"{max(r[3].batch_queries for r in reqs)}", flush=True)
    q_dev = torch.from_numpy(qpool).cuda()
    for q0, nq, k, res in reqs:
        assert res.k_eff == min(k, res.n_snapshot)
        assert res.t_submit <= res.t_start <= res.t_done
        s, i = ix.search_prefix_device(q_dev[q0:q0 + nq], k, res.n_snapshot)
        s, i = s.cpu().numpy(), i.cpu().numpy().view(np.uint64)
        assert bits_equal(s, res.results.scores), (q0, nq, k, res.n_snapshot, res.batch)
        assert np.array_equal(i, res.results.ids), (q0, nq, k, res.n_snapshot, res.batch)
    # adds finished in order, each reporting the size after it
    assert [a.n_snapshot for a in adds] == [n0 + 5000 * (j + 1) for j in range(8)]


def test_requests_against_oracle(served, oracle):
    import ctypes

    ix, qpool, xa, n0, reqs, adds, stats = served
    flat = ix.export_flat()
    alpha = ix.export_alpha()
    sh, sc = ix.calibration()
    picked = {}
    for q0, nq, k, res in reqs:
        picked.setdefault(res.n_snapshot, (q0, nq, k, res))
    for n_s, (q0, nq, k, res) in list(picked.items())[:3]:
        G = flat.shape[1]
        nb = (n_s + 191) // 192
        perm0 = np.zeros(nb * G * 192, np.uint8)
        f = np.ascontiguousarray(flat[:n_s])
        oracle._port("pack_perm0")(f.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n_s), ctypes.c_size_t(G),
                                   perm0.ctypes.data_as(ctypes.c_void_p))
        s, i = oracle.search_flat(qpool[q0:q0 + nq], k, 4, perm0, alpha[:n_s], sh, sc)
        assert bits_equal(s, res.results.scores) and np.array_equal(i, res.results.ids), (n_s, q0, nq, k)


def test_admission_refusals(vc):
    import torch

    from vectorcache_amd._capi import DeviceOutOfMemory, IndexStateError
    from vectorcache_amd.serving import Server, ServerParams, batch_bytes

    dim = 128
    rng = np.random.default_rng(5301)
    x = rng.standard_normal((3000, dim)).astype(np.float32)
    ix = vc.TurboQuantIndex(dim, 4)
    ix.add(x[:2000])
    ix.prepare()
    need = batch_bytes(ix, 3000, 64, 10)
    with pytest.raises(ValueError):  # one batch of max_batch_queries at max_k must fit the pool
        Server(ix, ServerParams(max_batch_queries=64, max_k=10, hbm_budget_bytes=need // 2))
    sv = Server(ix, ServerParams(slots=2, max_batch_queries=64, max_k=10, hbm_budget_bytes=4 * need,
                                 reserve_vectors=2500, max_vectors=2500, add_chunk=256))
    try:
        with pytest.raises(ValueError):
            sv.submit_search(x[:4], 11)  # k above max_k
        with pytest.raises(ValueError):
            sv.submit_search(x[:65], 10)  # more queries than a batch holds
        with pytest.raises(IndexStateError):
            ix.search(x[:2], 5)  # the index refuses direct use while served
        with pytest.raises(IndexStateError):
            ix.add(x[:2])
        with pytest.raises(IndexStateError):
            ix.close()
        t_ok = sv.submit_add(x[2000:2400])
        t_big = sv.submit_add(x[2400:3000])  # its first 256-row chunk passes max_vectors: refused
        t_q = sv.submit_search(x[:3], 10)
        assert sv.wait(t_ok).n_snapshot == 2400
        with pytest.raises(DeviceOutOfMemory):
            sv.wait(t_big)
        r = sv.wait(t_q)
        assert r.results.ids.shape == (3, 10)
        with pytest.raises(ValueError):
            sv.wait(t_q)  # a ticket is released by its wait
    finally:
        sv.close()
    assert ix.size() == 2400
    s, i = ix.search_prefix_device(torch.from_numpy(x[:3]).cuda(), 10, 2000)
    assert i.shape == (3, 10)
