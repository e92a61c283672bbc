"""Benchmarks of the MI355X TurboQuant path on one GPU.

Default workload `bucketed10m` (the config the headline metric is quoted on):
the bucketed index over 10,000,000 x 768 float32 rows from a seeded mixture
of 20,000 centers on the unit sphere with Zipf(1.1) cluster sizes and noise
of RMS norm 0.2, 4-bit codes, max_bucket_size 800, energy_soft 400; the
learned probe model trained on 180k queries (labels from the plain flat
scan); 10k held-out queries, k = 10. `value` is the learned-probe search's
queries/s at the recall@10 target (against the plain flat scan) that fixed
n_probe <= 256 reaches on validation queries; beside it the fixed-n_probe
QPS at the same target, the bucket scan's HBM roofline, latency and the
reference's CPU search_flat on a sample.

Other workloads: flat1m (1M x 1536 flat index, TQ+, 10k queries), flat1m_k64,
flat1m_b2, flat100k, landmark1m, landmark1m_planes (the same index searched
by the staged sign-plane branch), bucketed1m.

    python bench.py [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "QPS @ recall@10≥0.95 (learned vs fixed probe); p99 ms; % HBM roofline"

WORKLOADS = {
    # name: (n, dim, bits, calibrate_rows, nq, k)
    "flat1m": (1_000_000, 1536, 4, 20_000, 10_000, 10),
    "flat1m_k64": (1_000_000, 1536, 4, 20_000, 10_000, 64),
    "flat1m_b2": (1_000_000, 1536, 2, 20_000, 10_000, 10),
    "flat100k": (100_000, 128, 4, 0, 1_000, 10),
}

LANDMARK_WORKLOADS = {
    # name: (n, dim, scan_bits, n_landmarks, nq, k) -- the reference's LandmarkIndex
    # (TurboQuantIndex) with its default 1024 landmarks
    "landmark1m": (1_000_000, 1536, 4, 1024, 10_000, 10),
}
# The same index with sign planes kept: the staged search the reference runs
# on an AVX-512 VNNI host (sign-plane shortlist, then the permute-dot rescore).
LANDMARK_PLANES = {"landmark1m_planes": "landmark1m"}


SERVE_WORKLOADS = {
    # name: (n, dim, bits, calibrate_rows, add_rate_vectors_per_s, offered QPS levels, seconds per level)
    "serve1m": (1_000_000, 1536, 4, 20_000, 10_000,
                (5_000, 10_000, 20_000, 25_000, 30_000, 35_000, 40_000, 80_000, 160_000, 200_000), 3.0),
}


MULTI_WORKLOADS = {
    # name: (collections, n range, dims, bits, query batch range, ks, centers, sigma)
    "multi64": (64, (10_000, 2_000_000), (128, 768, 1536), (2, 3, 4), (1, 4096), (1, 10, 100), 512, 0.3),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def gen_data(name, n, dim, nq, cal_rows, device):
    import torch

    from vectorcache_amd import synthetic

    if name == "flat100k":
        x = torch.from_numpy(synthetic.flat100k_db(n)).to(device)
        q = torch.from_numpy(synthetic.flat100k_queries(nq)).to(device)
        return x, q, None
    x = synthetic.anisotropic_torch(n, dim, 2001, device)
    cal = synthetic.anisotropic_torch(cal_rows, dim, 2002, device) if cal_rows else None
    q = synthetic.anisotropic_torch(nq, dim, 2003, device)
    return x, q, cal


def exact_topk(x, q, k, chunk=200_000):
    """FP64 exact inner-product top-k (ties -> lower id) of normalized queries."""
    import torch

    qn = q.double()
    qn = qn / qn.norm(dim=1, keepdim=True)
    best_s = torch.full((q.shape[0], k), -math.inf, dtype=torch.float64, device=q.device)
    best_i = torch.zeros((q.shape[0], k), dtype=torch.int64, device=q.device)
    for s in range(0, x.shape[0], chunk):
        xs = x[s:s + chunk].double()
        sc = qn @ xs.T
        ids = torch.arange(s, s + xs.shape[0], device=q.device).expand(q.shape[0], -1)
        cs = torch.cat([best_s, sc], 1)
        ci = torch.cat([best_i, ids], 1)
        # stable sort on id then score keeps lower ids first among equal scores
        o1 = torch.argsort(ci, dim=1, stable=True)
        cs, ci = torch.gather(cs, 1, o1), torch.gather(ci, 1, o1)
        o2 = torch.argsort(-cs, dim=1, stable=True)[:, :k]
        best_s, best_i = torch.gather(cs, 1, o2), torch.gather(ci, 1, o2)
    return best_s, best_i


# ---- CPU baseline protocol ---------------------------------------------------
# The reference's OpenMP CPU path (oracle/_ref when built, else the C
# restatement) timed on this host: every physical core the process may use
# (OMP_PROC_BIND=close, OMP_PLACES=cores) and an 8-thread cell (the
# reference's bench convention, Makefile:31-34), each cell 2 warmup runs and
# the mean of 21 timed runs (tools/query-bench.cpp:526-607), repeated in
# fresh processes. A run is one search_flat over a query sample, or
# encode + codes_flat + repack_perm0 over a row sample; the sample is sized so
# that a run takes about CPU_RUN_S, which bounds the whole leg to ~1-2 min.
CPU_WARMUP, CPU_RUNS, CPU_REPEATS, CPU_RUN_S = 2, 21, 3, 0.25


def host_cpu_info():
    """lscpu's model and topology, the physical cores among the CPUs this
    process may run on, and the cgroup CPU quota when one is set."""
    info = {"model": cpu_model()}
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info["lscpu " + key.strip()] = val.strip()
        if info.get("lscpu Model name"):
            info["model"] = info["lscpu Model name"]
    except (OSError, ValueError):
        pass
    cpus = sorted(os.sched_getaffinity(0))
    phys = set()
    for c in cpus:
        topo = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            phys.add((topo.joinpath("physical_package_id").read_text().strip(),
                      topo.joinpath("core_id").read_text().strip()))
        except OSError:
            phys.add(("cpu", c))
    info["logical_cpus_allowed"] = len(cpus)
    info["physical_cores_allowed"] = len(phys)
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if quota != "max":
            info["cgroup_cpu_quota"] = int(quota) / int(period)
    except (OSError, ValueError):
        pass
    return info


def cpu_thread_cells(info):
    cores = max(1, info["physical_cores_allowed"])
    return [cores] if cores == 8 else [cores, 8]


def cpu_cell_main(spec_path):
    """One cell of the CPU baseline, in a child process started by
    run_cpu_cells: loads the shared arrays (memory-mapped), sizes the sample
    to about CPU_RUN_S per run unless the spec fixes it, then CPU_WARMUP
    untimed and CPU_RUNS timed runs; prints one JSON line and saves the last
    run's outputs beside the inputs."""
    spec = json.loads(Path(spec_path).read_text())
    d = Path(spec["dir"])
    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle import Oracle

    o = Oracle(spec["kind"])
    o.set_threads(spec["threads"])
    bits, dim = spec["bits"], spec["dim"]

    def ld(name):
        p = d / f"{name}.npy"
        return np.load(p, mmap_mode="r") if p.exists() else None

    sh, sc = ld("shift"), ld("scale")
    if spec["op"] == "search":
        q, perm0, alpha = ld("q"), ld("perm0"), ld("alpha")
        avail, step = q.shape[0], 1

        def run(m):
            return o.search_flat(q[:m], spec["k"], bits, perm0, alpha, sh, sc)
    else:
        x = ld("x")
        avail, step = x.shape[0] // 192 * 192, 192

        def run(m):
            packed, a = o.encode(x[:m], bits, sh, sc)
            return o.codes_flat(packed, bits, dim), o.repack_perm0(packed, bits, dim), a

    m = spec.get("m")
    if not m:
        # the CPU path splits a run over queries (search) or rows (encode): probe with every thread busy
        m0 = min(avail, 2 * spec["threads"] if spec["op"] == "search" else 4 * step)
        run(m0)  # pages the mapped inputs in
        t0 = time.perf_counter()
        run(m0)
        per = (time.perf_counter() - t0) / m0
        m = int(min(avail, max(step, CPU_RUN_S / max(per, 1e-9) // step * step)))
    for _ in range(CPU_WARMUP):
        out = run(m)
    ts = []
    for _ in range(CPU_RUNS):
        t0 = time.perf_counter()
        out = run(m)
        ts.append(time.perf_counter() - t0)
    names = ("scores", "ids") if spec["op"] == "search" else ("flat", "perm0", "alpha")
    for name, a in zip(names, out):
        np.save(d / f"out_{name}.npy", a)
    print(json.dumps({"m": m, "mean_s": float(np.mean(ts)), "min_s": float(np.min(ts)), "max_s": float(np.max(ts))}))


def run_cpu_cells(op, arrays, params, info):
    """Runs `op` ("search" or "encode") for every thread cell, CPU_REPEATS
    fresh processes each, over arrays written once to a temporary folder.
    Returns the cells (rate = sample / mean run time, mean over the repeats)
    and the last run's outputs."""
    import shutil
    import subprocess
    import tempfile

    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle import available

    kind = "reference" if available("reference") else "port"
    tmp = Path(tempfile.mkdtemp(prefix="vc_cpu_"))
    try:
        for name, a in arrays.items():
            if a is not None:
                np.save(tmp / f"{name}.npy", np.ascontiguousarray(a))
        cells = []
        for threads in cpu_thread_cells(info):
            m, reps = None, []
            for _ in range(CPU_REPEATS):
                spec = dict(params, op=op, kind=kind, threads=threads, dir=str(tmp), m=m)
                (tmp / "spec.json").write_text(json.dumps(spec))
                env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", OMP_NUM_THREADS=str(threads))
                p = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-cell", str(tmp / "spec.json")],
                                   env=env, capture_output=True, text=True, timeout=600)
                if p.returncode != 0:
                    raise RuntimeError(f"cpu cell ({op}, {threads} threads) failed: {p.stderr[-1500:]}")
                r = json.loads(p.stdout.strip().splitlines()[-1])
                m = r["m"]
                reps.append(r["mean_s"])
            cells.append({"threads": threads, "sample": m, "mean_s_per_run": reps,
                          "rate": m / float(np.mean(reps))})
        names = ("scores", "ids") if op == "search" else ("flat", "perm0", "alpha")
        outs = {nm: np.load(tmp / f"out_{nm}.npy") for nm in names}
        return kind, cells, outs
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_record(kind, info, search_cells, search_unit, search_sample, encode_cells=None, encode_sample=None, **extra):
    """The bench's cpu_baseline object: value / cores from the all-cores
    search cell, every cell beside it, the host description."""
    top = search_cells[0]
    rec = {"value": top["rate"], "unit": search_unit, "cores": top["threads"], "kind": kind,
           "cpu": info["model"], "host": info,
           "protocol": f"{CPU_WARMUP} warmups + mean of {CPU_RUNS} runs, {CPU_REPEATS} fresh processes per cell, "
                       "OMP_PROC_BIND=close OMP_PLACES=cores",
           "sample": search_sample, "search_cells": search_cells}
    if encode_cells is not None:
        rec["encode"] = {"unit": "vectors/s", "value": encode_cells[0]["rate"], "cores": encode_cells[0]["threads"],
                         "sample": encode_sample, "cells": encode_cells}
    rec.update(extra)
    return rec


def score_error_fp64(x, q, scores, ids):
    """Error of the quantized scores against the FP64 exact inner product they
    estimate, <q / |q|, x_id> (queries are normalized, rows are not), over the
    returned (query, id) pairs."""
    import torch

    qd = q.double()
    qd = qd / qd.norm(dim=1, keepdim=True).clamp_min(1e-30)
    idx = ids.long()
    ok = (idx >= 0) & (idx < x.shape[0])
    rows = x[idx.clamp(0, x.shape[0] - 1).view(-1)].double().view(idx.shape[0], idx.shape[1], -1)
    exact = torch.einsum("qd,qkd->qk", qd, rows)
    err = (scores.double() - exact)[ok]
    ex = exact[ok]
    return {"pairs": int(ok.sum()), "mean_abs": float(err.abs().mean()), "rms": float(err.pow(2).mean().sqrt()),
            "max_abs": float(err.abs().max()), "mean_signed": float(err.mean()),
            "mean_abs_exact": float(ex.abs().mean()),
            "relative_mean_abs": float(err.abs().mean() / ex.abs().mean().clamp_min(1e-30))}


def cpu_baseline(ix, x_rows, q_np, gpu_ids, gpu_scores, bits, dim, shift, scale):
    """The reference's own search_flat on a bounded query sample, over the GPU
    index's exported PERM0 codes and alpha (bit-identical to the reference's
    own encode), and its encode + codes_flat + repack_perm0 on a row sample;
    both checked against the GPU's results on the same sample."""
    info = host_cpu_info()
    k = gpu_ids.shape[1]
    perm0 = ix.export_perm0()
    prm = {"bits": bits, "dim": dim, "k": k}
    kind, s_cells, s_out = run_cpu_cells("search", {"q": q_np, "perm0": perm0, "alpha": ix.export_alpha(),
                                                    "shift": shift, "scale": scale}, prm, info)
    m = s_cells[0]["sample"]
    ms = s_cells[-1]["sample"]  # the saved outputs are the last cell's
    _, e_cells, e_out = run_cpu_cells("encode", {"x": x_rows, "shift": shift, "scale": scale}, prm, info)
    me = e_cells[-1]["sample"]
    G = dim // (8 // bits)
    return cpu_record(
        kind, info, s_cells, "queries/s",
        f"first {m} of the {q_np.shape[0]} queries per run (8-thread cell: {s_cells[-1]['sample']}), same index "
        "(search_flat over the exported PERM0 codes)",
        e_cells, f"first rows of the database per run ({', '.join(str(c['sample']) for c in e_cells)} rows), "
                 "encode + codes_flat + repack_perm0",
        parity_ids=bool(np.array_equal(s_out["ids"], gpu_ids[:ms])),
        parity_scores=bool(np.array_equal(s_out["scores"], gpu_scores[:ms])),
        encode_parity_perm0=bool(np.array_equal(e_out["perm0"], perm0[: me * G])))


def host_buffer_qps(ix, q, k, reps=2):
    """Queries per second through the host-buffer boundary (search(span)):
    query upload, search and result download included; never the reported
    value, which keeps inputs resident in HBM."""
    qh = q.cpu().numpy()
    ix.search(qh, k)
    t0 = time.perf_counter()
    for _ in range(reps):
        ix.search(qh, k)
    return qh.shape[0] * reps / (time.perf_counter() - t0)


def copy_probe(dev, nbytes=1 << 30, reps=5):
    """Device-to-device copy of a known size: the achievable HBM ceiling
    (read + write) and the reference size for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "gb_per_s": 2 * nbytes / ms / 1e6}


def committed_traffic(workload, nq, tile):
    """HBM bytes per launch of the dominant scan launch, from the committed
    counter profile of this workload (FETCH_SIZE and WRITE_SIZE, each scaled
    by what the counter reads for the copy probe's known size); None when
    no profile of this workload and query tile is committed."""
    p = ROOT / "profiles" / f"scan_pair_kernel_{workload}_traffic.json"
    if not p.exists():
        return None, None
    d = json.loads(p.read_text())
    if d.get("nq") != nq or d.get("query_tile") != tile:
        return None, None
    return d["traffic_bytes_per_launch"], str(p.relative_to(ROOT))


BUCKETED_WORKLOADS = {
    # name: (n, dim, bits, n_centers, zipf_s, sigma, n_train, n_val, n_eval, k, max_bucket_size, energy_soft)
    "bucketed10m": (10_000_000, 768, 4, 20_000, 1.1, 0.2, 180_000, 20_000, 10_000, 10, 800, 400),
    "bucketed1m": (1_000_000, 768, 4, 2_000, 1.1, 0.2, 45_000, 5_000, 5_000, 10, 800, 400),
}
RECALL_TARGETS = (0.95, 0.9, 0.85, 0.8, 0.75, 0.7, 0.6, 0.5)
# Lookup ceiling of 8-query tables: conflict-free 64-bit LDS reads measured at 3.3 CU-cycles per wavefront
# (64 lanes x 8 queries), 256 CUs at 2.4 GHz.
LDS_LOOKUP_CEILING = 64 * 8 / 3.3 * 256 * 2.4e9
LDS_LOOKUP_SOURCE = "conflict-free 8-byte LDS reads, 3.3 CU-cycles per wavefront (64 lanes x 8 queries), 256 CUs, 2.4 GHz"


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def chunked(n, size):
    for s in range(0, n, size):
        yield s, min(n, s + size)


def probe_ranks(ix, q, judge_ids, bucket_of_row, P, chunk=10_000):
    """Rank, in each query's routing order, of the bucket holding each of its
    judge top-k rows (P = not among its P best buckets)."""
    import torch

    out = torch.empty(judge_ids.shape, dtype=torch.int32, device=q.device)
    for s, e in chunked(q.shape[0], chunk):
        pb, _ = ix.route_device(q[s:e], P)
        bo = bucket_of_row[judge_ids[s:e]]
        hit = pb.unsqueeze(1) == bo.unsqueeze(2)
        r = torch.argmax(hit.to(torch.int8), dim=2).to(torch.int32)
        out[s:e] = torch.where(hit.any(dim=2), r, torch.full_like(r, P))
    return out


def recall_at(ranks, n_probe):
    """Mean recall@k against the judge when query i opens n_probe[i] buckets."""
    import torch

    n = n_probe if torch.is_tensor(n_probe) else torch.full((ranks.shape[0],), int(n_probe), device=ranks.device)
    return float((ranks < n.view(-1, 1)).float().mean())


PROBE_GRID = tuple(range(1, 17)) + tuple(range(20, 65, 4)) + tuple(range(72, 257, 8))


def fit_probe_table(y, ranks, csz, target, y0=0.0, y_step=0.25, n_bins=40, grid=PROBE_GRID):
    """Probe counts per bin of the model's prediction y (bin = floor((y - y0) /
    y_step), clamped) that reach a mean recall of `target` with the fewest
    scanned vectors on these queries: each step moves the bin whose next
    larger count (any later grid point) gains the most recall per scanned
    vector; the last move is the cheapest one that reaches the target.
    ranks: routing rank of each judge row's bucket; csz: cumulative
    vectors of each query's routing list."""
    import torch

    g = torch.tensor(grid, device=ranks.device)
    g = g[g <= csz.shape[1]]
    rec = (ranks.unsqueeze(2) < g.view(1, 1, -1)).float().mean(1)  # [n, G]
    cost = csz[:, g - 1].double()                                   # [n, G]
    b = torch.clamp(torch.floor((y - y0) / y_step), 0, n_bins - 1).long()
    R = torch.zeros((n_bins, g.numel()), dtype=torch.float64, device=ranks.device).index_add_(0, b, rec.double())
    C = torch.zeros((n_bins, g.numel()), dtype=torch.float64, device=ranks.device).index_add_(0, b, cost)
    R, C, gl = R.cpu().numpy(), C.cpu().numpy(), g.cpu().numpy()
    n = ranks.shape[0]
    at = np.zeros(n_bins, np.int64)
    total = R[np.arange(n_bins), at].sum()
    cols = np.arange(len(gl))
    while total < target * n:
        dr = R - R[np.arange(n_bins), at][:, None]
        dc = C - C[np.arange(n_bins), at][:, None]
        ok = (cols[None, :] > at[:, None]) & (dr > 0) & (dc > 0)
        if not ok.any():
            break
        reach = ok & (total + dr >= target * n)
        if reach.any():
            # the last move: the cheapest one that reaches the target (not the best ratio, which can overshoot)
            flat = int(np.argmin(np.where(reach, dc, np.inf)))
        else:
            ratio = np.where(ok, dr / np.where(dc > 0, dc, 1.0), -1.0)
            flat = int(np.argmax(ratio))  # first of the best (bin, count), as the scan in bin, count order
        pick = divmod(flat, len(gl))
        total += R[pick[0], pick[1]] - R[pick[0], at[pick[0]]]
        at[pick[0]] = pick[1]
    return gl[at].astype(np.int32)


def oracle_policy(ranks, csz, target):
    """The fewest scanned vectors any per-query probe count reaches `target`
    mean recall with, on these queries: an oracle that knows each query's
    judge rows opens, per query, just enough of its routing list (at most
    csz.shape[1] buckets) to take in some of them. Chosen by a Lagrange
    multiplier on recall against vectors (each query takes the count that
    maximizes recall - lambda * vectors), bisected until the mean recall meets
    the target: a point of the trade-off's lower convex hull, within one
    query's vectors of the cheapest combination of per-query counts, so a
    bound the learned policy can be held against. ranks: routing rank of each
    judge row's bucket; csz: cumulative vectors of each query's routing list."""
    import torch

    cap = csz.shape[1]
    n, k = ranks.shape
    r, _ = torch.sort(ranks.long(), dim=1)
    reach = r < cap
    p = torch.clamp(r + 1, max=cap)                                   # probes to take in rows 0 .. j
    rec = (ranks.long().unsqueeze(1) < p.unsqueeze(2)).double().mean(2)  # [n, k]
    cost = torch.gather(csz, 1, p - 1).double()
    cost0 = csz[:, 0].double()
    rows = torch.arange(n, device=ranks.device)

    def pick(lam):
        val = torch.where(reach, rec - lam * cost, torch.full_like(cost, -np.inf))
        best, j = val.max(1)
        take = best > -lam * cost0
        recall = torch.where(take, rec[rows, j], torch.zeros_like(best))
        vecs = torch.where(take, cost[rows, j], cost0)
        probes = torch.where(take, p[rows, j], torch.ones_like(j))
        return float(recall.mean()), float(vecs.sum()), float(probes.double().mean())

    lo, hi = 0.0, 1.0
    while pick(hi)[0] >= target and hi < 1e6:
        hi *= 2.0
    if pick(lo)[0] < target:
        return None
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if pick(mid)[0] >= target:
            lo = mid
        else:
            hi = mid
    recall, vecs, probes = pick(lo)
    return {"recall": recall, "vectors": vecs, "mean_n_probe": probes}


def table_counts(y, table, valid, y0=0.0, y_step=0.25):
    import torch

    t = torch.from_numpy(table).to(y.device)
    b = torch.clamp(torch.floor((y - y0) / y_step), 0, len(table) - 1).long()
    return torch.minimum(t[b], valid.to(torch.int32)).to(torch.int32)


def run_bucketed(args):
    """The bucketed index with the learned probe budget: QPS at a recall@10
    target against the plain flat scan, learned vs fixed n_probe."""
    import torch

    from vectorcache_amd import synthetic
    from vectorcache_amd.bucketed import BucketedTurboQuantIndex, BucketParams
    import vectorcache_amd.index as vc

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n, dim, bits, nc, zs, sigma, n_tr, n_val, n_ev, k, mbs, es = BUCKETED_WORKLOADS[args.workload]
    G = dim // (8 // bits)
    log(f"workload {args.workload}: n={n} dim={dim} bits={bits} centers={nc} zipf={zs} sigma={sigma} "
        f"train={n_tr} val={n_val} eval={n_ev} k={k}")
    t0 = time.perf_counter()
    x = synthetic.zipf_mixture_torch(n, dim, nc, zs, sigma, 4001, 4002, dev)
    q_tr = synthetic.zipf_mixture_torch(n_tr + n_val, dim, nc, zs, sigma, 4001, 4003, dev)
    q_ev = synthetic.zipf_mixture_torch(n_ev, dim, nc, zs, sigma, 4001, 4004, dev)
    torch.cuda.synchronize()
    log(f"data generated in {time.perf_counter() - t0:.1f} s")

    ix = BucketedTurboQuantIndex(dim, bits, BucketParams(n_probe=64, max_bucket_size=mbs, energy_soft=es,
                                                         residual=False, seed=4001))
    ix.add(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix.prepare()
    build_s = time.perf_counter() - t0
    L = ix.n_buckets()
    st = ix.export()
    sizes = torch.from_numpy(st["bucket_sizes"].astype(np.int64)).to(dev)
    pos_ids = torch.from_numpy(st["pos_ids"].astype(np.int64)).to(dev)
    bucket_of_row = torch.empty(n, dtype=torch.int32, device=dev)
    bucket_of_row[pos_ids] = torch.repeat_interleave(torch.arange(L, device=dev, dtype=torch.int32), sizes)
    log(f"build {build_s:.1f} s ({n / build_s / 1e6:.2f} M vectors/s): {L} buckets, sizes "
        f"{int(sizes.min())}..{int(sizes.max())} mean {float(sizes.float().mean()):.0f}")

    # ---- the judge: the plain flat scan over every code, then routing ranks ----
    P = min(512, L)
    t0 = time.perf_counter()
    q_all = torch.cat([q_tr, q_ev])
    judge = torch.empty((q_all.shape[0], k), dtype=torch.int64, device=dev)
    for s, e in chunked(q_all.shape[0], 10_000):
        _, i = ix.search_exhaustive_device(q_all[s:e], k)
        judge[s:e] = i
    torch.cuda.synchronize()
    judge_s = time.perf_counter() - t0
    log(f"judge: flat scan of {q_all.shape[0]} queries in {judge_s:.1f} s ({q_all.shape[0] / judge_s:.0f} QPS)")
    ranks = probe_ranks(ix, q_all, judge, bucket_of_row, P)
    labels = ranks.max(dim=1).values + 1  # buckets needed for the whole judge top-k (P + 1: not within P)
    feats = torch.cat([ix.probe_features(q_all[s:e]) for s, e in chunked(q_all.shape[0], 10_000)])
    n_tv = n_tr + n_val
    tr, va, ev = slice(0, n_tr), slice(n_tr, n_tv), slice(n_tv, n_tv + n_ev)
    lab_np = labels.cpu().numpy()
    log("labels (buckets for recall 1.0) quantiles 10/50/90/99: "
        f"{np.percentile(lab_np[tr], [10, 50, 90, 99]).tolist()}")

    # ---- the recall target: the highest of RECALL_TARGETS that fixed n_probe <= 256 reaches on validation ----
    fixed_curve = {p: recall_at(ranks[va], p) for p in range(1, min(256, L) + 1)}
    target = next((t for t in RECALL_TARGETS if max(fixed_curve.values()) >= t), RECALL_TARGETS[-1])
    # labels at that target: per query the smallest n_probe whose result reaches recall@k >= target
    # against the judge (for 0.95 and k = 10: every judge row, the largest rank + 1)
    q_lab = min(k, int(math.ceil(target * k - 1e-9))) - 1
    labels_t = torch.sort(ranks, dim=1).values[:, q_lab] + 1
    log(f"labels at recall {target} (rank of judge row {q_lab + 1} of {k}) quantiles 10/50/90/99: "
        f"{np.percentile(labels_t[tr].cpu().numpy(), [10, 50, 90, 99]).tolist()}")

    # ---- train the probe model on the GPU ----
    targets = torch.log2(torch.clamp(labels_t, 1, 512).float())
    ix.probe_init(4005)
    rng = np.random.default_rng(4006)
    epochs = 8
    t0 = time.perf_counter()
    loss_first = loss_last = None
    for ep in range(epochs):
        order = torch.from_numpy(rng.permutation(n_tr).astype(np.int32)[: n_tr // 64 * 64]).to(dev)
        losses = ix.probe_train(feats[tr], targets[tr], order, lr=1e-3 if ep < 6 else 3e-4)
        lc = losses.cpu().numpy()
        loss_first = float(lc[:50].mean()) if loss_first is None else loss_first
        loss_last = float(lc[-200:].mean())
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"probe model: {epochs} epochs of {n_tr} queries in {train_s:.1f} s, loss {loss_first:.3f} -> {loss_last:.3f}")

    # ---- recall target, fixed n_probe and the learned policy, chosen on the validation split ----
    valid = torch.clamp(torch.full((q_all.shape[0],), min(256, L), device=dev), max=P)
    y = ix.probe_predict(feats)
    n_fixed = next(p for p, r in fixed_curve.items() if r >= target) if max(fixed_curve.values()) >= target else min(
        256, L)
    # the learned policy: the model's prediction picks a bin, each bin's probe count is fitted on validation
    csz_all = torch.cat([torch.cumsum(sizes[ix.route_device(q_all[s:e], min(256, L))[0].long()], dim=1)
                         for s, e in chunked(q_all.shape[0], 10_000)])
    # the learned table is fitted to the fixed policy's validation recall (at least the target), so that the
    # two are compared at the same recall
    table = fit_probe_table(y[va], ranks[va], csz_all[va], max(target, fixed_curve[n_fixed]))
    ix.probe_set_table(table, 0.0, 0.25)
    ix.probe_enable(True, 0.0)
    cnt_ev = table_counts(y[ev], table, valid[ev])
    rec_fixed_ev = recall_at(ranks[ev], n_fixed)
    rec_learned_ev = recall_at(ranks[ev], cnt_ev)
    reach = {p: recall_at(ranks[ev], p) for p in (1, 2, 4, 8, 16, 32, 64, 128, 256) if p <= L}
    log(f"recall target {target}: fixed n_probe {n_fixed} (eval recall {rec_fixed_ev:.4f}), learned table {table.tolist()} "
        f"mean n_probe {float(cnt_ev.float().mean()):.1f} (eval recall {rec_learned_ev:.4f}); curve {reach}")

    # ---- timed: the eval queries, fixed vs learned ----
    qe = q_ev.contiguous()
    stream = torch.cuda.current_stream().cuda_stream
    vc.set_profiling(True)
    res = {}
    for mode, n_probe in (("fixed", n_fixed), ("learned", 0)):
        sc = torch.empty((n_ev, k), dtype=torch.float32, device=dev)
        ids = torch.empty((n_ev, k), dtype=torch.int64, device=dev)
        for _ in range(args.warmup):
            ix.search_device(qe, k, n_probe=n_probe, scores=sc, ids=ids, stream=stream)
        torch.cuda.synchronize()
        scan_ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ix.search_device(qe, k, n_probe=n_probe, scores=sc, ids=ids, stream=stream)
            scan_ms.append(vc.last_scan_stats()[0])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        scan_tile = vc.last_scan_stats()[1]
        got = ids.cpu().numpy()
        jd = judge[ev].cpu().numpy()
        rec = float(np.mean([len(set(got[i]) & set(jd[i])) / k for i in range(n_ev)]))
        res[mode] = {"qps": n_ev * args.steps / el, "ms_per_step": el / args.steps * 1e3,
                     "scan_ms": float(np.mean(scan_ms)), "scan_tile": scan_tile, "recall_at_10_vs_flat_scan": rec,
                     "ids": ids, "scores": sc}
        log(f"{mode}: {res[mode]['qps']:.0f} QPS, {res[mode]['ms_per_step']:.2f} ms/step, bucket scan "
            f"{res[mode]['scan_ms']:.2f} ms, recall@10 vs flat scan {rec:.4f}")

    # vectors the bucket scan reads per pass (algorithmic bytes: codes + alpha of the opened buckets, LUT rows)
    csz = csz_all[ev]
    vec_fixed = int(csz[:, n_fixed - 1].sum())
    vec_learned = int(torch.gather(csz, 1, (cnt_ev.long() - 1).view(-1, 1)).sum())
    oracle = oracle_policy(ranks[ev], csz, rec_fixed_ev)
    if oracle is not None:
        oracle["vectors_fixed_over_oracle"] = vec_fixed / oracle["vectors"]
        oracle["vectors_fixed_over_learned"] = vec_fixed / max(vec_learned, 1)
        log(f"oracle probe policy at recall {rec_fixed_ev:.4f}: {oracle}")
    # Algorithmic bytes of one scan launch. Query-major (tile 1): every query reads its opened buckets' codes
    # and alpha and its LUT rows. Bucket-major (tile 8): every tile of up to 8 queries of one bucket reads the
    # bucket once and its 8 LUT rows. Query-nibble lookups: 2 per code byte per (query, vector) pair.
    pb_ev = torch.cat([ix.route_device(qe[s_:e_], min(256, L))[0] for s_, e_ in chunked(n_ev, 10_000)]).long()
    scan_bytes, tile_info = {}, {}
    for mode, cnt in (("fixed", torch.full((n_ev,), n_fixed, device=dev)), ("learned", cnt_ev.long())):
        vec = vec_fixed if mode == "fixed" else vec_learned
        tile = res[mode]["scan_tile"]
        if tile > 1:
            mask = torch.arange(pb_ev.shape[1], device=dev)[None, :] < cnt[:, None]
            qb = torch.bincount(pb_ev[mask], minlength=L)
            tiles = (qb + tile - 1) // tile
            t_vec = int((tiles * sizes).sum())
            scan_bytes[mode] = t_vec * (G + 4) + int(tiles.sum()) * tile * G * 32
            tile_info[mode] = {"tiles": int(tiles.sum()), "tile_vectors": t_vec, "pairs": int(qb.sum()),
                               "query_slot_use": vec / max(t_vec * tile, 1)}
        else:
            scan_bytes[mode] = vec * (G + 4) + n_ev * G * 32
    bytes_l, bytes_f = scan_bytes["learned"], scan_bytes["fixed"]
    ach_l = bytes_l / (res["learned"]["scan_ms"] / 1e3) / 1e9
    ach_f = bytes_f / (res["fixed"]["scan_ms"] / 1e3) / 1e9
    lookups_l = vec_learned * 2 * G / (res["learned"]["scan_ms"] / 1e3)

    # ---- quality against FP64 exact, latency ----
    ns = min(200, n_ev)
    es_, ei = exact_topk(x, qe[:ns], k)
    got = res["learned"]["ids"][:ns].cpu().numpy()
    rec64 = float(np.mean([len(set(got[i]) & set(ei.cpu().numpy()[i])) / k for i in range(ns)]))
    got_j = judge[ev][:ns].cpu().numpy()
    rec64_flat = float(np.mean([len(set(got_j[i]) & set(ei.cpu().numpy()[i])) / k for i in range(ns)]))
    score_err = score_error_fp64(x, qe[:ns], res["learned"]["scores"][:ns], res["learned"]["ids"][:ns])
    log(f"score error vs fp64 (learned, {ns} queries): {score_err}")
    lat = {}
    if not args.no_latency:
        for b in (1, 64):
            ts = []
            qb = qe[:b].contiguous()
            for it in range(25):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                ix.search_device(qb, k, stream=stream)
                torch.cuda.synchronize()
                if it >= 5:
                    ts.append((time.perf_counter() - t1) * 1e3)
            lat[f"batch{b}"] = {"p50_ms": float(np.percentile(ts, 50)), "p99_ms": float(np.percentile(ts, 99))}
        log(f"latency (learned) {lat}")

    probe = copy_probe(dev)
    log(f"copy probe {probe}")
    # HBM traffic per launch from the committed counter run of the same kernel and workload; the learned
    # policy's table comes out of a training run, so its opened buckets (and bytes) vary slightly between runs:
    # the committed figure is used when its run's algorithmic bytes are within 2 % of this run's
    traffic, traffic_src, traffic_alg = None, None, None
    kname = "bucket_tile_kernel" if res["learned"]["scan_tile"] > 1 else "bucket_scan_kernel"
    tp = ROOT / "profiles" / f"{kname}_{args.workload}_traffic.json"
    if tp.exists():
        td = json.loads(tp.read_text())
        ta = td.get("learned_algorithmic_bytes") or 0
        if td.get("nq") == n_ev and td.get("n_buckets") == L and abs(ta - bytes_l) <= 0.02 * bytes_l:
            traffic, traffic_src, traffic_alg = (td["learned_traffic_bytes_per_launch"], str(tp.relative_to(ROOT)),
                                                 ta)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = bucketed_cpu_baseline(st, x[:65_472].cpu().numpy(), q_ev[:1024].cpu().numpy(),
                                        judge[ev][:1024].cpu().numpy(), bits, dim, k, n)
            log(f"cpu baseline {cpu}")
        except Exception as e:  # reported, never fatal for the GPU figure
            cpu = {"error": str(e)}

    out = {
        "metric": METRIC,
        "value": res["learned"]["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["learned"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded, generated on device)",
        "config": {
            "workload": f"{args.workload}: bucketed TurboQuant index, {n} x {dim} float32 from {nc} centers on the "
                        f"unit sphere, Zipf({zs}) cluster sizes, noise RMS {sigma}, {bits}-bit codes of the rows, "
                        f"max_bucket_size {mbs}, energy_soft {es}; probe MLP 96-256-256-1 trained on {n_tr} queries; "
                        f"{n_ev} held-out queries, k={k}",
            "n": n, "dim": dim, "bits": bits, "k": k, "n_buckets": L, "nq": n_ev,
            "recall_target": target, "recall_judge": "the plain flat scan over every code (exact LUT top-10)",
            "n_probe_fixed": n_fixed, "n_probe_learned_mean": float(cnt_ev.float().mean()),
            "learned_probe_table": {"y0": 0.0, "y_step": 0.25, "counts": table.tolist()},
            "seeds": {"centers": 4001, "rows": 4002, "train_val_queries": 4003, "eval_queries": 4004},
        },
        "fixed": {**{k2: v for k2, v in res["fixed"].items() if k2 not in ("ids", "scores")},
                  "tiles": tile_info.get("fixed"), "algorithmic_bytes": bytes_f},
        "learned": {k2: v for k2, v in res["learned"].items() if k2 not in ("ids", "scores")},
        "learned_over_fixed_qps": res["learned"]["qps"] / res["fixed"]["qps"],
        "eval_recall_from_ranks": {"fixed": rec_fixed_ev, "learned": rec_learned_ev},
        "oracle_probe_policy": oracle,
        "recall_curve_fixed_n_probe": reach,
        "label_quantiles_10_50_90_99": np.percentile(lab_np[ev], [10, 50, 90, 99]).tolist(),
        "training_labels": {"recall": target, "judge_row": q_lab + 1,
                            "quantiles_10_50_90_99": np.percentile(labels_t[ev].cpu().numpy(),
                                                                   [10, 50, 90, 99]).tolist()},
        "roofline": {
            "bound": "hbm", "achieved": ach_l, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach_l / HBM_PEAK_GBS,
            "traffic": traffic, "traffic_source": traffic_src, "traffic_run_algorithmic_bytes": traffic_alg,
            "kernel": ("bucket_tile_kernel" if res["learned"]["scan_tile"] > 1 else "bucket_scan_kernel")
                      + " (learned probe)",
            "query_tile": res["learned"]["scan_tile"],
            "avg_ms": res["learned"]["scan_ms"],
            "algorithmic_bytes": bytes_l,
            "tiles": tile_info.get("learned"),
            "lookups": {"achieved": lookups_l, "ceiling": LDS_LOOKUP_CEILING, "unit": "query-nibble lookups/s",
                        "frac": lookups_l / LDS_LOOKUP_CEILING, "ceiling_source": LDS_LOOKUP_SOURCE},
            "note": ("bytes = per tile of up to 8 queries of one bucket, the bucket's codes and alpha once + the "
                     "tile's 8 LUT rows" if res["learned"]["scan_tile"] > 1 else
                     "bytes = vectors of the opened buckets x (G+4) + LUT rows")
                    + f"; {vec_learned} query-vector pairs learned, {vec_fixed} fixed (fixed: {ach_f:.0f} GB/s)",
        },
        "build": {"seconds": build_s, "vectors_per_s": n / build_s, "judge_seconds": judge_s,
                  "judge_qps": q_all.shape[0] / judge_s, "train_seconds": train_s, "loss_first": loss_first,
                  "loss_last": loss_last},
        "recall_at_10_vs_fp64": {"learned": rec64, "flat_scan": rec64_flat, "queries": ns},
        "score_error_vs_fp64": score_err,
        "latency": lat,
        "copy_probe": probe,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def bucketed_cpu_baseline(st, x_rows, q_np, judge_ids, bits, dim, k, n_full):
    """The reference's own search_flat over all the index's codes (PERM0
    blocks of the exported flat codes) on a bounded query sample: the CPU's
    exhaustive answer to the same question (the reference has no GPU bucketed
    path to time), checked against the GPU flat scan; and the rate of its
    encode + codes_flat + repack_perm0 on a row sample."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle import Oracle

    info = host_cpu_info()
    flat = st["flat"]
    G = flat.shape[1]
    nb = (flat.shape[0] + 191) // 192
    perm0 = np.zeros(nb * G * 192, np.uint8)
    Oracle("port")._port("pack_perm0")(flat.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(flat.shape[0]),
                                       ctypes.c_size_t(G), perm0.ctypes.data_as(ctypes.c_void_p))
    prm = {"bits": bits, "dim": dim, "k": k}
    kind, s_cells, s_out = run_cpu_cells("search", {"q": q_np, "perm0": perm0, "alpha": st["alpha"]}, prm, info)
    del perm0
    ms = s_cells[-1]["sample"]  # the saved outputs are the last cell's
    ids = st["pos_ids"][s_out["ids"].astype(np.int64)]
    # the rate of the CPU encode of raw rows; the index encodes the rows scaled by their inverse norms, so its
    # codes are not this encode's output (the flat leg checks encode parity)
    _, e_cells, _ = run_cpu_cells("encode", {"x": x_rows}, prm, info)
    return cpu_record(
        kind, info, s_cells, "queries/s",
        f"first {s_cells[0]['sample']} held-out queries per run (8-thread cell: {ms}), exhaustive search_flat "
        f"over all {n_full} codes (the reference has no bucketed index)",
        e_cells, f"first rows of the database per run ({', '.join(str(c['sample']) for c in e_cells)} rows), "
                 "encode + codes_flat + repack_perm0",
        parity_ids_vs_gpu_flat_scan=bool(np.array_equal(ids, judge_ids[:ms].astype(np.uint64))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="bucketed10m",
                    choices=sorted(WORKLOADS) + sorted(LANDMARK_WORKLOADS) + sorted(LANDMARK_PLANES)
                    + sorted(BUCKETED_WORKLOADS)
                    + sorted(SERVE_WORKLOADS) + sorted(MULTI_WORKLOADS))
    ap.add_argument("--tile", type=int, default=0, help="cap the scan query tile (1, 2, 4, 8)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-latency", action="store_true")
    ap.add_argument("--serve-slots", type=int, default=0, help="serve workloads: worker slots (streams)")
    ap.add_argument("--serve-batch", type=int, default=0, help="serve workloads: max queries per batch")
    ap.add_argument("--serve-levels", default="", help="serve workloads: comma-separated offered QPS levels")
    ap.add_argument("--cpu-cell", help=argparse.SUPPRESS)  # child process of the CPU baseline
    args = ap.parse_args()
    if args.cpu_cell:
        return cpu_cell_main(args.cpu_cell)
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the flat scan would shard by rows "
                                   "with an all-gather of per-GPU top-k lists", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload in LANDMARK_WORKLOADS or args.workload in LANDMARK_PLANES:
        return run_landmark(args)
    if args.workload in SERVE_WORKLOADS:
        return run_serve(args)
    if args.workload in MULTI_WORKLOADS:
        return run_multi(args)
    if args.workload in BUCKETED_WORKLOADS:
        return run_bucketed(args)

    import torch

    import vectorcache_amd.index as vc

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n, dim, bits, cal_rows, nq, k = WORKLOADS[args.workload]
    G = dim // (8 // bits)
    log(f"workload {args.workload}: n={n} dim={dim} bits={bits} nq={nq} k={k}")
    x, q, cal = gen_data(args.workload, n, dim, nq, cal_rows, dev)
    torch.cuda.synchronize()

    # ---- encode: time prepare() of the whole database (inputs resident) ----
    enc_times = []
    ix = None
    for rep in range(3):
        if ix is not None:
            ix.close()
        ix = vc.TurboQuantIndex(dim, bits)
        if cal is not None:
            ix.calibrate(cal)
        ix.add(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ix.prepare()
        enc_times.append(time.perf_counter() - t0)
    enc_s = min(enc_times)
    enc_bytes = n * (4 * dim + G + 4)
    log(f"encode {n / enc_s / 1e6:.1f} M vectors/s ({enc_bytes / enc_s / 1e9:.0f} GB/s)")
    shift, scale = ix.calibration() if cal is not None else (None, None)

    # ---- search ----
    if args.tile:
        vc.set_scan_tile(args.tile)
    stream = torch.cuda.current_stream().cuda_stream
    vc.set_profiling(True)
    scores = torch.empty((nq, k), dtype=torch.float32, device=dev)
    ids = torch.empty((nq, k), dtype=torch.int64, device=dev)
    for _ in range(args.warmup):
        ix.search_device(q, k, scores, ids, stream=stream)
    torch.cuda.synchronize()
    scan_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ix.search_device(q, k, scores, ids, stream=stream)
        ms, tile, ranges = vc.last_scan_stats()
        scan_ms.append(ms)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    qps = nq * args.steps / elapsed
    log(f"search {qps:.0f} QPS, {elapsed / args.steps * 1e3:.1f} ms/step, scan kernel {np.mean(scan_ms):.1f} ms "
        f"(tile {tile}, ranges {ranges})")

    passes = math.ceil(nq / tile)
    scan_bytes = passes * n * (G + 4) + nq * G * 32
    scan_avg_s = float(np.mean(scan_ms)) / 1e3
    achieved = scan_bytes / scan_avg_s / 1e9
    lookups = nq * n * 2 * G / scan_avg_s

    # ---- bandwidth regime: small query batches, scan kernel alone ----
    bw = []
    for nqb in (1, 8):
        qb = q[:nqb].contiguous()
        sb = torch.empty((nqb, k), dtype=torch.float32, device=dev)
        ib = torch.empty((nqb, k), dtype=torch.int64, device=dev)
        ms_b = []
        for it in range(12):
            ix.search_device(qb, k, sb, ib, stream=stream)
            ms_i, tile_b, ranges_b = vc.last_scan_stats()
            if it >= 2:
                ms_b.append(ms_i)
        t_b = float(np.median(ms_b)) / 1e3
        bytes_b = math.ceil(nqb / tile_b) * n * (G + 4) + nqb * G * 32
        bw.append({"nq": nqb, "query_tile": tile_b, "vector_ranges": ranges_b, "scan_ms": t_b * 1e3,
                   "achieved_gbs": bytes_b / t_b / 1e9, "frac": bytes_b / t_b / 1e9 / HBM_PEAK_GBS})
    log(f"bandwidth regime {bw}")

    # ---- quality: recall@10 vs FP64 exact and the scores' error on a query sample ----
    ns = min(200, nq)
    es, ei = exact_topk(x, q[:ns], 10)
    got = ids[:ns, :10].cpu().numpy()
    exact = ei.cpu().numpy()
    recall = float(np.mean([len(set(got[i]) & set(exact[i])) / 10 for i in range(ns)]))
    score_err = score_error_fp64(x, q[:ns], scores[:ns], ids[:ns])
    log(f"recall@10 vs fp64 exact ({ns} queries): {recall:.4f}; score error vs fp64 {score_err}")

    # ---- latency of small request batches ----
    lat = {}
    if not args.no_latency:
        for b in (1, 64):
            ts = []
            qb = q[:b].contiguous()
            sb = torch.empty((b, k), dtype=torch.float32, device=dev)
            ib = torch.empty((b, k), dtype=torch.int64, device=dev)
            for it in range(25):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                ix.search_device(qb, k, sb, ib, stream=stream)
                torch.cuda.synchronize()
                if it >= 5:
                    ts.append((time.perf_counter() - t1) * 1e3)
            lat[f"batch{b}"] = {"p50_ms": float(np.percentile(ts, 50)), "p99_ms": float(np.percentile(ts, 99))}
        log(f"latency {lat}")

    probe = copy_probe(dev)
    log(f"copy probe {probe}")
    host_qps = host_buffer_qps(ix, q, k)
    log(f"host-buffer search {host_qps:.0f} QPS (PCIe-inclusive)")

    # ---- CPU baseline (reference build) ----
    cpu = None
    if not args.no_cpu:
        try:
            x_rows = x[: min(n, 65_472)].cpu().numpy()
            cpu = cpu_baseline(ix, x_rows, q.cpu().numpy(), ids.cpu().numpy().astype(np.uint64),
                               scores.cpu().numpy(), bits, dim, shift, scale)
            log(f"cpu baseline {cpu}")
        except Exception as e:  # reported, never fatal for the GPU figure
            cpu = {"error": str(e)}

    traffic, traffic_src = committed_traffic(args.workload, nq, tile)
    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded, generated on device)",
        "config": {
            "workload": f"{args.workload}: plain flat TurboQuant, {n} x {dim} float32, {bits}-bit"
                        + (f", TQ+ on {cal_rows} rows" if cal_rows else "") + f", {nq} queries, k={k}",
            "n": n, "dim": dim, "bits": bits, "nq": nq, "k": k, "query_tile": tile, "vector_ranges": ranges,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "kernel": f"scan_pair_kernel (query tile {tile})", "avg_ms": scan_avg_s * 1e3,
            "algorithmic_bytes": scan_bytes,
            "note": f"bytes = ceil(nq/T) passes x N x (G+4) + LUTs; T={tile}; nibble lookups/s = {lookups:.3e}",
        },
        "encode": {"vectors_per_s": n / enc_s, "ms": enc_s * 1e3, "gb_per_s": enc_bytes / enc_s / 1e9,
                   "frac": enc_bytes / enc_s / 1e9 / HBM_PEAK_GBS},
        "roofline_bandwidth_regime": bw,
        "host_buffer_qps": host_qps,
        "copy_probe": probe,
        "recall_at_10_vs_fp64": recall,
        "score_error_vs_fp64": score_err,
        "latency": lat,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def _host_has_vnni():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return all(f in flags for f in ("avx512_vnni", "avx512vbmi", "avx512bw"))


def landmark_cpu_baseline(x_np, q_np, bits, m, k, n_full, budget_queries=1000, n_sample=100_000, planes=False):
    """The reference's own LandmarkIndex (oracle/_ref) built on the first
    n_sample vectors, its search() of budget_queries queries timed; the rate is
    scaled to the full index (the scan is linear in n). Also checks a GPU
    index of the same sample against it. planes: the reference's VNNI build
    (its staged search), which needs an AVX-512 VNNI host."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle import Oracle, available

    from vectorcache_amd.landmark import LandmarkIndex

    kind = "reference" if available("reference") else "port"
    if planes and (kind != "reference" or not _host_has_vnni()):
        return {"error": "the staged search's CPU figure needs the reference build on an AVX-512 VNNI host"}
    o = Oracle(kind)
    host = host_cpu_info()
    cores = host["physical_cores_allowed"]
    o.set_threads(cores)
    xs = np.ascontiguousarray(x_np[:n_sample])
    qs = np.ascontiguousarray(q_np[:budget_queries])
    s, i, info = o.lm_search(xs, qs, k, scan_bits=bits, n_landmarks=m, seed=0, planes=planes)
    dt = info["search_s"]
    g = LandmarkIndex(xs.shape[1], bits, m, planes=planes)
    g.add(xs)
    g.prepare()
    r = g.search(qs, k)
    g.close()
    out = {
        "unit": "queries/s", "cores": cores, "kind": kind, "cpu": host["model"], "host": host,
        "sample": f"LandmarkIndex on the first {n_sample} of the {n_full} vectors, search() of {len(qs)} queries; "
                  f"rate x {n_sample}/{n_full} (scan linear in n)",
        "parity_ids": bool(np.array_equal(r.ids, i)), "parity_scores": bool(np.array_equal(r.scores, s)),
    }
    if dt > 0:
        out["value"] = len(qs) / dt * n_sample / n_full
        out["sample_queries_per_s"] = len(qs) / dt
    return out


def serve_level(sv, qpool, rate_qps, seconds, rng, max_nq=64, ks=(10, 100)):
    """Open-loop Poisson arrivals of requests (1..max_nq queries, k drawn from
    ks) at rate_qps queries/s for `seconds`; latency is measured from each
    request's scheduled arrival (late submits count against the server)."""
    import threading

    mean_nq = (1 + max_nq) / 2
    n_req = max(1, int(rate_qps / mean_nq * seconds))
    gaps = rng.exponential(mean_nq / rate_qps, n_req)
    nqs = rng.integers(1, max_nq + 1, n_req)
    kss = rng.choice(ks, n_req)
    q0s = rng.integers(0, qpool.shape[0] - max_nq, n_req)
    t0 = time.monotonic() + 0.02
    due = t0 + np.cumsum(gaps)
    tickets = []
    done = []

    def collect():
        i = 0
        while i < n_req:
            if i >= len(tickets):
                time.sleep(0.0002)
                continue
            t, a, nq, k = tickets[i]
            r = sv.wait(t)
            done.append((a, nq, k, r.t_start, r.t_done, r.batch_queries, r.n_snapshot))
            i += 1

    col = threading.Thread(target=collect)
    col.start()
    late = 0.0
    for j in range(n_req):
        now = time.monotonic()
        if due[j] > now:
            time.sleep(max(0.0, due[j] - now - 0.0002))
            while time.monotonic() < due[j]:
                pass
        else:
            late = max(late, now - due[j])
        nq = int(nqs[j])
        q0 = int(q0s[j])
        tickets.append((sv.submit_search(qpool[q0:q0 + nq], int(kss[j])), float(due[j]), nq, int(kss[j])))
    col.join()
    a = np.array([d[0] for d in done])
    td = np.array([d[4] for d in done])
    lat = (td - a) * 1e3
    nq_tot = int(sum(d[1] for d in done))
    span = float(td.max() - t0)
    return {"offered_qps": rate_qps, "requests": n_req, "queries": nq_tot, "seconds": seconds,
            "achieved_qps": nq_tot / span, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "p999_ms": float(np.percentile(lat, 99.9)),
            "mean_batch_queries": float(np.mean([d[5] for d in done])),
            "queue_ms_p50": float(np.percentile((np.array([d[3] for d in done]) - a) * 1e3, 50)),
            "max_submit_lag_ms": late * 1e3,
            "index_size_first_last": [int(min(d[6] for d in done)), int(max(d[6] for d in done))]}


def run_serve(args):
    """The request server over the 1M x 1536 4-bit TQ+ flat index while it
    grows by a concurrent add stream: open-loop Poisson arrivals at rising
    offered rates until the server saturates; p50 / p99 latency beside the
    achieved QPS at each rate."""
    import threading

    import torch

    import vectorcache_amd.index as vc
    from vectorcache_amd import synthetic
    from vectorcache_amd.serving import Server, ServerParams

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n, dim, bits, cal_rows, add_rate, levels, secs = SERVE_WORKLOADS[args.workload]
    if args.serve_levels:
        levels = tuple(int(v) for v in args.serve_levels.split(","))
    G = dim // (8 // bits)
    log(f"workload {args.workload}: n={n} dim={dim} bits={bits} adds {add_rate}/s levels {levels}")
    x = synthetic.anisotropic_torch(n, dim, 2001, dev)
    cal = synthetic.anisotropic_torch(cal_rows, dim, 2002, dev)
    qpool = synthetic.anisotropic_torch(16_384, dim, 5001, dev).cpu().numpy()
    n_add_max = int(add_rate * secs * (len(levels) + 2) * 2)
    adds = synthetic.anisotropic_torch(n_add_max, dim, 5002, dev).cpu().numpy()
    ix = vc.TurboQuantIndex(dim, bits)
    ix.calibrate(cal)
    ix.add(x)
    ix.prepare()
    torch.cuda.synchronize()
    shift, scale = ix.calibration()
    params = ServerParams(slots=args.serve_slots or 4, max_batch_queries=args.serve_batch or 4096, max_k=100,
                          hbm_budget_bytes=2 << 30, reserve_vectors=n + n_add_max, add_chunk=8192)
    sv = Server(ix, params)
    stop = threading.Event()
    added = []

    def add_stream():  # add_rate vectors/s in tenths of a second
        pos, step, t_next = 0, add_rate // 10, time.monotonic()
        while not stop.is_set() and pos + step <= adds.shape[0]:
            t_next += 0.1
            added.append(sv.submit_add(adds[pos:pos + step]))
            pos += step
            time.sleep(max(0.0, t_next - time.monotonic()))

    rng = np.random.default_rng(5001)
    ta = threading.Thread(target=add_stream)
    ta.start()
    out_levels = []
    try:
        for rate in levels:
            lv = serve_level(sv, qpool, rate, secs, rng)
            lv["saturated"] = bool(lv["achieved_qps"] < 0.9 * rate or lv["p99_ms"] > 1000.0)
            out_levels.append(lv)
            log(f"offered {rate} QPS: achieved {lv['achieved_qps']:.0f}, p50 {lv['p50_ms']:.2f} ms, p99 "
                f"{lv['p99_ms']:.2f} ms, mean batch {lv['mean_batch_queries']:.0f} queries"
                + (" (saturated)" if lv["saturated"] else ""))
            if lv["saturated"]:
                break
    finally:
        stop.set()
        ta.join()
        add_infos = [sv.wait(t) for t in added]
        stats = sv.stats()
        sv.close()
    log(f"server stats {stats}; index grew {n} -> {ix.size()}")
    ok = [lv for lv in out_levels if not lv["saturated"]]
    best = max(ok, key=lambda lv: lv["achieved_qps"]) if ok else out_levels[0]

    # the scan kernel of a served-size batch alone (HIP events on its stream): the roofline line
    stream = torch.cuda.current_stream().cuda_stream
    vc.set_profiling(True)
    nb = 512
    qb = torch.from_numpy(qpool[:nb]).to(dev)
    sb = torch.empty((nb, 100), dtype=torch.float32, device=dev)
    ib = torch.empty((nb, 100), dtype=torch.int64, device=dev)
    ms_b = []
    for it in range(8):
        ix.search_device(qb, 10, sb, ib, stream=stream)
        ms_i, tile_b, ranges_b = vc.last_scan_stats()
        if it >= 2:
            ms_b.append(ms_i)
    torch.cuda.synchronize()
    t_b = float(np.median(ms_b)) / 1e3
    n_now = ix.size()
    bytes_b = math.ceil(nb / tile_b) * n_now * (G + 4) + nb * G * 32
    ach = bytes_b / t_b / 1e9

    cpu = None
    if not args.no_cpu:
        try:
            qh = qpool[:1024]
            s_g, i_g = ix.search_device(torch.from_numpy(qh).to(dev), 10)
            cpu = cpu_baseline(ix, x[:65_472].cpu().numpy(), qh,
                               i_g.cpu().numpy().astype(np.uint64), s_g.cpu().numpy(), bits, dim, shift, scale)
            log(f"cpu baseline {cpu}")
        except Exception as e:  # reported, never fatal for the GPU figure
            cpu = {"error": str(e)}
    out = {
        "metric": METRIC,
        "value": best["achieved_qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": len(out_levels),
        "warmup": 0,
        "ms_per_step": secs * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded, generated on device)",
        "config": {
            "workload": f"{args.workload}: request server over the flat TurboQuant index, {n} x {dim} float32, "
                        f"{bits}-bit, TQ+ on {cal_rows} rows, growing by {add_rate} vectors/s; open-loop Poisson "
                        f"arrivals, requests of 1-64 queries, k in {{10, 100}}",
            "n": n, "dim": dim, "bits": bits, "levels_qps": list(levels), "seconds_per_level": secs,
            "server": params.__dict__, "seeds": {"arrivals": 5001, "adds": 5002},
        },
        "value_is": "achieved QPS at the highest offered rate the server sustained (achieved >= 0.9 offered, "
                    "p99 < 1 s)",
        "latency": {"p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"], "at_offered_qps": best["offered_qps"]},
        "levels": out_levels,
        "server_stats": stats,
        "adds": {"requests": len(add_infos), "vectors": int(sum(a.k_eff for a in add_infos)),
                 "p50_ms": float(np.percentile([(a.t_done - a.t_submit) * 1e3 for a in add_infos], 50))
                 if add_infos else None},
        "roofline": {
            "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "kernel": f"scan_pair_kernel (a {nb}-query batch alone, query tile {tile_b})",
            "avg_ms": t_b * 1e3, "algorithmic_bytes": bytes_b,
            "note": f"bytes = ceil(nq/T) passes x N x (G+4) + LUTs over the grown index (N={n_now})",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def multi_specs(name):
    """The collections of a multi workload: per collection j (seed 3000 + j)
    its size log-uniform in the range, dim, bits, query batch log-uniform in
    the range, and k."""
    m, (n_lo, n_hi), dims, bitss, (b_lo, b_hi), ks, nc, sigma = MULTI_WORKLOADS[name]
    out = []
    for j in range(m):
        r = np.random.default_rng(3000 + j)
        n = int(round(math.exp(r.uniform(math.log(n_lo), math.log(n_hi)))))
        dim = int(r.choice(dims))
        bits = int(r.choice(bitss))
        nq = int(round(math.exp(r.uniform(math.log(b_lo), math.log(b_hi)))))
        k = int(r.choice(ks))
        out.append((n, dim, bits, nq, k))
    return out, nc, sigma


def run_multi(args):
    """Many collections of uneven size, dim and bits, each with its own query
    batch and k: the whole set searched by one persistent scan launch (graph
    replay of the prepare -> scan -> merge chain, autotuned plan) against the
    same collections searched one after another; every collection's results
    checked bit for bit against its own search."""
    import torch

    import vectorcache_amd.index as vc
    from vectorcache_amd import synthetic
    from vectorcache_amd.multi import MultiSearch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    specs, nc, sigma = multi_specs(args.workload)
    t0 = time.perf_counter()
    idx, qs = [], []
    for j, (n, dim, bits, nq, k) in enumerate(specs):
        x = synthetic.zipf_mixture_torch(n, dim, nc, 0.0, sigma, 3000 + j, 3100 + j, dev)
        ix = vc.TurboQuantIndex(dim, bits)
        ix.add(x)
        ix.prepare()
        del x
        idx.append(ix)
        qs.append(synthetic.zipf_mixture_torch(nq, dim, nc, 0.0, sigma, 3000 + j, 3200 + j, dev))
    torch.cuda.synchronize()
    n_tot = sum(s[0] for s in specs)
    nq_tot = sum(s[3] for s in specs)
    log(f"workload {args.workload}: {len(specs)} collections, {n_tot} vectors, {nq_tot} queries; built in "
        f"{time.perf_counter() - t0:.1f} s")
    stream = torch.cuda.current_stream().cuda_stream

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t1) / steps

    # the collections one after another (each its own search_device)
    outs = [None] * len(specs)

    def sequential():
        for j, (ix, q) in enumerate(zip(idx, qs)):
            outs[j] = ix.search_device(q, specs[j][4], stream=stream)

    seq_s = timed(sequential, args.steps)
    want = [(s.cpu().numpy(), i.cpu().numpy()) for s, i in outs]
    log(f"sequential: {seq_s * 1e3:.1f} ms per pass ({nq_tot / seq_s:.0f} QPS)")

    ms = MultiSearch(idx, [s[3] for s in specs], [s[4] for s in specs])
    for j, q in enumerate(qs):
        ms.set_queries(j, q)
    default_plan = ms.plan()
    eager_default_s = timed(lambda: ms.search(stream=stream), args.steps)
    log(f"one launch, default plan {default_plan}: {eager_default_s * 1e3:.1f} ms ({nq_tot / eager_default_s:.0f} QPS)")
    t1 = time.perf_counter()
    tuned = ms.tune(reps=2)
    tune_s = time.perf_counter() - t1
    eager_s = timed(lambda: ms.search(stream=stream), args.steps)
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    ms.capture(side.cuda_stream)
    graph_s = timed(lambda: ms.replay(stream=stream), args.steps)
    log(f"one launch, tuned plan {tuned} (tuned in {tune_s:.1f} s): eager {eager_s * 1e3:.1f} ms, graph replay "
        f"{graph_s * 1e3:.1f} ms ({nq_tot / graph_s:.0f} QPS)")
    parity = True
    for j in range(len(specs)):
        s, i = ms.results(j)
        if not (np.array_equal(s.cpu().numpy().view(np.uint32), want[j][0].view(np.uint32))
                and np.array_equal(i.cpu().numpy(), want[j][1])):
            parity = False
            log(f"collection {j} {specs[j]} differs from its own search")
    # the scan launch alone, HIP events on its stream
    vc.set_profiling(True)
    scan_ms = []
    for it in range(6):
        ms.search(stream=stream)
        m_i, tile, items = vc.last_scan_stats()
        if it >= 1:
            scan_ms.append(m_i)
    vc.set_profiling(False)
    t_scan = float(np.median(scan_ms)) / 1e3
    tiles = [ms.collection_plan(j)[0] for j in range(len(specs))]
    scan_bytes = sum(math.ceil(nq / tiles[j]) * n * (dim * bits // 8 + 4) + nq * (dim * bits // 8) * 32
                     for j, (n, dim, bits, nq, k) in enumerate(specs))
    ach = scan_bytes / t_scan / 1e9
    lookups = sum(nq * n * 2 * (dim * bits // 8) for n, dim, bits, nq, k in specs) / t_scan
    log(f"scan launch {t_scan * 1e3:.2f} ms: {ach:.0f} GB/s algorithmic ({ach / HBM_PEAK_GBS:.2f} of HBM)")
    ms.close()

    cpu = None
    if not args.no_cpu:
        try:
            # the reference's search_flat over one collection of median cost (its own queries)
            cost = [math.ceil(nq / 8) * n * dim * bits for n, dim, bits, nq, k in specs]
            j = int(np.argsort(cost)[len(cost) // 2])
            n, dim, bits, nq, k = specs[j]
            qh = qs[j].cpu().numpy()
            info = host_cpu_info()
            kind, cells, outc = run_cpu_cells("search", {"q": qh, "perm0": idx[j].export_perm0(),
                                                         "alpha": idx[j].export_alpha()},
                                              {"bits": bits, "dim": dim, "k": k}, info)
            m_s = cells[-1]["sample"]
            cpu = cpu_record(kind, info, cells, "queries/s",
                             f"collection {j} (n={n}, dim={dim}, bits={bits}, k={k}; median cost of the set), the "
                             f"first {cells[0]['sample']} of its {nq} queries per run",
                             parity_ids=bool(np.array_equal(outc["ids"], want[j][1][:m_s].astype(np.uint64))))
            log(f"cpu baseline {cpu}")
        except Exception as e:  # reported, never fatal for the GPU figure
            cpu = {"error": str(e)}
    out = {
        "metric": METRIC,
        "value": nq_tot / graph_s,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": graph_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded, generated on device)",
        "config": {
            "workload": f"{args.workload}: {len(specs)} flat TurboQuant collections, sizes log-uniform in "
                        f"[10k, 2M] ({n_tot} vectors), dim in {{128, 768, 1536}}, bits in {{2, 3, 4}}, each from "
                        f"its own {nc}-center mixture (sigma {sigma}); query batches log-uniform in [1, 4096] "
                        f"({nq_tot} queries), k in {{1, 10, 100}}; one persistent scan launch, graph replay",
            "collections": [{"n": n, "dim": dim, "bits": bits, "nq": nq, "k": k} for n, dim, bits, nq, k in specs],
        },
        "sequential": {"ms": seq_s * 1e3, "qps": nq_tot / seq_s},
        "one_launch_default_plan": {"ms": eager_default_s * 1e3, "plan": default_plan.__dict__},
        "one_launch_tuned": {"eager_ms": eager_s * 1e3, "graph_ms": graph_s * 1e3, "plan": tuned.__dict__,
                             "tune_seconds": tune_s},
        "speedup_vs_sequential": seq_s / graph_s,
        "parity_all_collections_bit_exact": parity,
        "roofline": {
            "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "kernel": f"multi_scan_kernel (tiles up to {tuned.tile}, {items} work items)",
            "avg_ms": t_scan * 1e3, "algorithmic_bytes": scan_bytes,
            "note": f"bytes = sum over collections of ceil(nq/T_c) passes x n x (G+4) + LUTs, T_c the collection's "
                    f"tile; nibble lookups/s = {lookups:.3e}",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def run_landmark(args):
    """The landmark index (the reference's TurboQuantIndex): add + prepare of
    the whole database, then search of the query batch (device-resident)."""
    import torch

    import vectorcache_amd.index as vc
    from vectorcache_amd import synthetic
    from vectorcache_amd.landmark import LandmarkIndex

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    planes = args.workload in LANDMARK_PLANES
    n, dim, bits, m, nq, k = LANDMARK_WORKLOADS[LANDMARK_PLANES.get(args.workload, args.workload)]
    S = (m + 7) // 8 * 8
    G = S // (8 // bits)
    log(f"workload {args.workload}: n={n} dim={dim} scan_bits={bits} landmarks={m} nq={nq} k={k}")
    x = synthetic.anisotropic_torch(n, dim, 2001, dev)
    q = synthetic.anisotropic_torch(nq, dim, 2003, dev)
    LandmarkIndex(dim, bits, m).close()  # the basis (host Gram-Schmidt) is memoized per process
    ix = LandmarkIndex(dim, bits, m, planes=planes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix.add(x)
    t_add = time.perf_counter() - t0
    t0 = time.perf_counter()
    ix.prepare()
    t_prep = time.perf_counter() - t0
    log(f"add {n / t_add / 1e6:.2f} M vectors/s, prepare {n / t_prep / 1e6:.2f} M vectors/s")
    if planes and not ix.planes_active():
        raise RuntimeError("sign planes were not kept for the staged workload")
    # the scanned bytes per vector: the sign plane (S / 8) in the staged search
    G_scan = S // 8 if planes else G

    if args.tile:
        vc.set_scan_tile(args.tile)
    stream = torch.cuda.current_stream().cuda_stream
    vc.set_profiling(True)
    scores = torch.empty((nq, k), dtype=torch.float32, device=dev)
    ids = torch.empty((nq, k), dtype=torch.int64, device=dev)
    for _ in range(args.warmup):
        ix.search_device(q, k, scores, ids, stream=stream)
    torch.cuda.synchronize()
    scan_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ix.search_device(q, k, scores, ids, stream=stream)
        ms, tile, ranges = vc.last_scan_stats()
        scan_ms.append(ms)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    qps = nq * args.steps / elapsed
    scan_avg_s = float(np.mean(scan_ms)) / 1e3
    scan_bytes = math.ceil(nq / tile) * n * (G_scan + 4) + nq * G_scan * 32
    achieved = scan_bytes / scan_avg_s / 1e9
    log(f"search {qps:.0f} QPS, {elapsed / args.steps * 1e3:.1f} ms/step, scan kernel {scan_avg_s * 1e3:.1f} ms "
        f"(tile {tile}, ranges {ranges})")

    ns = min(200, nq)
    es, ei = exact_topk(x, q[:ns], 10)
    got = ids[:ns, :10].cpu().numpy()
    exact = ei.cpu().numpy()
    recall = float(np.mean([len(set(got[i]) & set(exact[i])) / 10 for i in range(ns)]))
    log(f"recall@10 vs fp64 exact ({ns} queries): {recall:.4f}")

    lat = {}
    if not args.no_latency:
        for b in (1, 64):
            ts = []
            qb = q[:b].contiguous()
            sb = torch.empty((b, k), dtype=torch.float32, device=dev)
            ib = torch.empty((b, k), dtype=torch.int64, device=dev)
            for it in range(25):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                ix.search_device(qb, k, sb, ib, stream=stream)
                torch.cuda.synchronize()
                if it >= 5:
                    ts.append((time.perf_counter() - t1) * 1e3)
            lat[f"batch{b}"] = {"p50_ms": float(np.percentile(ts, 50)), "p99_ms": float(np.percentile(ts, 99))}
        log(f"latency {lat}")

    host_qps = host_buffer_qps(ix, q, k)
    log(f"host-buffer search {host_qps:.0f} QPS (PCIe-inclusive)")
    probe = copy_probe(dev)
    log(f"copy probe {probe}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = landmark_cpu_baseline(x.cpu().numpy(), q.cpu().numpy(), bits, m, k, n, planes=planes)
            log(f"cpu baseline {cpu}")
        except Exception as e:  # reported, never fatal for the GPU figure
            cpu = {"error": str(e)}

    traffic, traffic_src = committed_traffic(args.workload, nq, tile)
    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded, generated on device)",
        "config": {
            "workload": f"{args.workload}: landmark index (LandmarkIndex), {n} x {dim} float32, {m} landmarks, "
                        f"{bits}-bit scan codes, {nq} queries, k={k}"
                        + (", sign planes (staged search of an AVX-512 VNNI host)" if planes else ""),
            "n": n, "dim": dim, "scan_bits": bits, "n_landmarks": m, "scan_dim": S, "nq": nq, "k": k,
            "planes": planes,
            "query_tile": tile, "vector_ranges": ranges,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "kernel": f"scan_pair_kernel (query tile {tile})", "avg_ms": scan_avg_s * 1e3,
            "algorithmic_bytes": scan_bytes,
            "note": f"bytes = ceil(nq/T) passes x N x (G+4) + LUTs; G={G_scan} "
                    + ("sign bytes of the landmark rows (the staged search's first step)" if planes
                       else "groups of the landmark rows") + f"; T={tile}",
        },
        "build": {"add_vectors_per_s": n / t_add, "prepare_vectors_per_s": n / t_prep,
                  "add_ms": t_add * 1e3, "prepare_ms": t_prep * 1e3},
        "recall_at_10_vs_fp64": recall,
        "latency": lat,
        "host_buffer_qps": host_qps,
        "copy_probe": probe,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
