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
