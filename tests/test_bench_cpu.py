"""The bench's CPU baseline leg on a small index: the fresh-process cells
(2 warmups + 21 timed runs per process) return the same top-k and the same
encode bytes as calling the CPU path directly, and the host description
names the cores it counted."""
import numpy as np

import bench


def test_host_cpu_info_counts_cores():
    info = bench.host_cpu_info()
    assert info["physical_cores_allowed"] >= 1
    assert info["logical_cpus_allowed"] >= info["physical_cores_allowed"]
    cells = bench.cpu_thread_cells(info)
    assert cells[0] == info["physical_cores_allowed"] and cells[-1] == 8


def test_cpu_cells_match_direct_calls(oracle, monkeypatch):
    monkeypatch.setattr(bench, "CPU_REPEATS", 1)
    monkeypatch.setattr(bench, "CPU_RUNS", 3)
    rng = np.random.default_rng(11)
    n, dim, bits, k = 3000, 128, 4, 10
    x = rng.standard_normal((n, dim)).astype(np.float32)
    q = rng.standard_normal((24, dim)).astype(np.float32)
    sh, sc = oracle.fit_calibration(x[:1000], bits)
    packed, alpha = oracle.encode(x, bits, sh, sc)
    perm0 = oracle.repack_perm0(packed, bits, dim)
    info = dict(bench.host_cpu_info(), physical_cores_allowed=2)
    prm = {"bits": bits, "dim": dim, "k": k}
    kind, cells, out = bench.run_cpu_cells("search", {"q": q, "perm0": perm0, "alpha": alpha, "shift": sh,
                                                      "scale": sc}, prm, info)
    assert [c["threads"] for c in cells] == [2, 8]
    m = cells[-1]["sample"]
    s, i = oracle.search_flat(q[:m], k, bits, perm0, alpha, sh, sc)
    assert np.array_equal(out["ids"], i) and np.array_equal(out["scores"], s)
    _, cells, out = bench.run_cpu_cells("encode", {"x": x[:960], "shift": sh, "scale": sc}, prm, info)
    me = cells[-1]["sample"]
    assert me % 192 == 0 and me <= 960
    assert np.array_equal(out["perm0"], perm0[: me * dim // 2])
    assert np.array_equal(out["alpha"], alpha[:me])
    assert all(c["rate"] > 0 for c in cells)


def test_oracle_probe_policy_against_brute_force():
    """On tiny instances every combination of per-query probe counts is
    tried: the oracle meets the target, and scans at most one query's
    vectors more than the cheapest combination that meets it."""
    import itertools

    import torch

    rng = np.random.default_rng(3)
    for trial in range(30):
        n, k, cap = 4, 3, 6
        ranks = torch.from_numpy(rng.integers(0, cap + 2, (n, k)).astype(np.int32))
        csz = torch.from_numpy(np.cumsum(rng.integers(1, 50, (n, cap)), axis=1))
        counts = list(range(1, cap + 1))
        rec = [[float((ranks[i] < p).float().mean()) for p in counts] for i in range(n)]
        cost = [[float(csz[i, p - 1]) for p in counts] for i in range(n)]
        best_rec = float(np.mean([max(r) for r in rec]))
        target = float(rng.uniform(0.1, 1.0)) * best_rec
        if target <= 0:
            continue
        best = min(sum(cost[i][c[i]] for i in range(n))
                   for c in itertools.product(range(cap), repeat=n)
                   if np.mean([rec[i][c[i]] for i in range(n)]) >= target - 1e-12)
        o = bench.oracle_policy(ranks, csz, target)
        assert o is not None and o["recall"] >= target - 1e-12, trial
        slack = max(max(c) - min(c) for c in cost)
        assert best - 1e-9 <= o["vectors"] <= best + slack, (trial, o, best)
