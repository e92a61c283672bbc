"""Writes the golden vectors under tests/golden/ from the reference's own code.

Run in a container that has the reference sources (it builds oracle/_ref from
them with oracle/Makefile and loads it); the outputs are data only:
inputs plus the reference's outputs for the plain flat TurboQuant path
(rotation tables, codebooks, encode, TQ+ calibration, flat bytes, PERM0
blocks, per-query LUTs, search_flat top-k, id_map scoring, result merging,
score_prepared_into over chained buckets, float_rescore_topk), run on the
reference's AVX2 PERM0 backend (TURBOVEC_NO_VNNI=1).

Large outputs are stored as SHA-256 digests of their little-endian bytes.

    python tests/golden/make_golden.py              # everything
    python tests/golden/make_golden.py score_into   # score_into.npz only
    python tests/golden/make_golden.py rescore      # float_rescore.npz only
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))

from oracle import Oracle  # noqa: E402
from vectorcache_amd import synthetic  # noqa: E402
from cases import (CASES, KS, LANDMARK_CASES, RESCORE_DIMS, SCORE_PREPARED_CASES, SCORE_PREPARED_VARIANTS,  # noqa: E402
                   case_data, landmark_inputs, rescore_inputs, score_prepared_inputs)

OUT = Path(__file__).resolve().parent


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _d(a: np.ndarray) -> np.ndarray:
    """SHA-256 digest of an array's bytes, as 32 uint8."""
    return np.frombuffer(bytes.fromhex(sha(a)), np.uint8)


ROT_DIMS = [8, 16, 24, 64, 128, 200, 256, 384, 768, 1024, 1536, 4096]
CB_DIMS = [8, 16, 64, 128, 200, 256, 512, 768, 1024, 1536, 3072, 16384]


# score_prepared_into scenarios: (name, bits, dim, tq, bucket sizes, ks)
SCORE_INTO = [("a", 4, 64, False, (300, 200, 300), (1, 7, 150)),
              ("b", 2, 128, True, (250, 97), (10,))]


def score_into_fixture(R):
    """Chained score_prepared_into calls per scenario: buckets of one dataset
    (with duplicated rows and zero rows, so scores tie and zero scores occur),
    each with its own id map, query subset (unordered) and score offsets
    (none, then a mix of 0, -0, positive and negative, then all +0); every
    heap starts empty at -inf except one query's, seeded at a finite
    threshold. Stores the inputs and every heap array after every call."""
    out = {}
    for name, bits, dim, tq, sizes, ks in SCORE_INTO:
        r = np.random.default_rng(4100 + dim)
        n = sum(sizes)
        centers = r.standard_normal((24, dim)).astype(np.float32)
        x = (centers[r.integers(0, 24, n)] + 0.3 * r.standard_normal((n, dim))).astype(np.float32)
        x[7] = 0.0
        x[n - 3] = 0.0
        x[20:26] = x[19]
        x[sizes[0] + 5: sizes[0] + 9] = x[3]
        nq = 6
        q = r.standard_normal((nq, dim)).astype(np.float32)
        shift = scale = None
        if tq:
            shift, scale = R.fit_calibration(x[:200], bits)
        packed, alpha = R.encode(x, bits, shift, scale)
        out[f"{name}_q"] = q
        out[f"{name}_bits"] = np.array([bits], np.int64)
        if tq:
            out[f"{name}_shift"] = shift
            out[f"{name}_scale"] = scale
        qsets = [np.array([3, 0, 5, 1], np.uint64), np.array([5, 2, 0, 4, 1, 3], np.uint64),
                 np.array([1, 5], np.uint64)]
        offs = [None, np.array([0.0, -0.0, 0.25, -1.5, 3e-3, 0.0], np.float32), np.zeros(nq, np.float32)]
        start = 0
        calls = []
        for ci, sz in enumerate(sizes):
            pk = packed[start:start + sz]
            perm0 = R.repack_perm0(pk, bits, dim)
            al = alpha[start:start + sz].copy()
            id_map = (np.arange(sz, dtype=np.uint64) * (ci + 2) + 1000 * (ci + 1)).astype(np.uint64)
            if ci == 2:
                id_map = id_map[::-1].copy()
            calls.append((perm0, al, id_map, qsets[ci % 3], offs[ci % 3]))
            out[f"{name}_c{ci}_perm0"] = perm0
            out[f"{name}_c{ci}_alpha"] = al
            out[f"{name}_c{ci}_idmap"] = id_map
            out[f"{name}_c{ci}_qidx"] = qsets[ci % 3]
            if offs[ci % 3] is not None:
                out[f"{name}_c{ci}_offsets"] = offs[ci % 3]
            start += sz
        for k in ks:
            heaps = (np.zeros((nq, k), np.float32), np.zeros((nq, k), np.uint64), np.zeros(nq, np.uint64),
                     np.full(nq, -np.inf, np.float32), np.zeros(nq, np.uint64))
            heaps[3][4] = np.float32(0.05)
            for ci, (perm0, al, id_map, qidx, off) in enumerate(calls):
                heaps = R.score_into(q, k, bits, perm0, al, qidx, heaps, id_map=id_map, shift=shift, scale=scale,
                                     offsets=off)
                for hn, h in zip(("hs", "hi", "hsz", "hmin", "hmi"), heaps):
                    out[f"{name}_k{k}_c{ci}_{hn}"] = h
            s, i = R.heaps_to_results(heaps[0], heaps[1], heaps[2])
            out[f"{name}_k{k}_scores"] = s
            out[f"{name}_k{k}_ids"] = i
    np.savez_compressed(OUT / "score_into.npz", **out)


def rescore_fixture(R):
    out = {}
    for dim in RESCORE_DIMS:
        corpus, q, scores, ids, counts = rescore_inputs(dim)
        for tag, cnt in (("all", None), ("counts", counts)):
            s, i = R.float_rescore(scores, ids, q, corpus, cnt)
            out[f"d{dim}_{tag}_scores"] = s
            out[f"d{dim}_{tag}_ids"] = i
    np.savez_compressed(OUT / "float_rescore.npz", **out)


def score_prepared_fixture(R):
    """score_prepared (AVX2 path) with id_map, query_indices and the residual
    post-correction, from the reference build."""
    out = {}
    for ci, (bits, dim, n, nq, k) in enumerate(SCORE_PREPARED_CASES):
        x, q, id_map, qi, rn, rs = score_prepared_inputs(bits, dim, n, nq)
        packed, alpha = R.encode(x, bits)
        perm0 = R.repack_perm0(packed, bits, dim)
        for tag, kw in SCORE_PREPARED_VARIANTS(id_map, qi, rn, rs):
            s, i = R.score_prepared(q, k, bits, perm0, alpha, **kw)
            out[f"c{ci}_{tag}_scores"] = s
            out[f"c{ci}_{tag}_ids"] = i
    np.savez_compressed(OUT / "score_prepared.npz", **out)


def landmark_fixture(R):
    """LandmarkIndex end to end (reference src/index.cpp:120-289) per case, the
    landmark basis and its 4-bit form for the small shapes, and the q4 query
    projection."""
    out = {}
    for tag, n, nq, dim, m, bits, rb, seed, k, batches, pb, fc in LANDMARK_CASES:
        x, q = landmark_inputs(n, nq, dim)
        s, i, info = R.lm_search(x, q, k, scan_bits=bits, n_landmarks=m, seed=seed, rerank_bits=rb,
                                 batches=batches, prepare_between=pb, float_corpus=x if fc else None)
        out[f"{tag}_scores"] = s
        out[f"{tag}_ids"] = i
        out[f"{tag}_info"] = np.array([info["scan_dim"], info["codes_bytes"], info["index_bytes"]], np.uint64)
    for dim, m, seed in ((64, 20, 0), (128, 100, 7)):
        L = R.ortho_landmarks(dim, m, seed)
        codes, scales = R.quantize_landmarks(L)
        if dim * m <= 4096:
            out[f"L_{dim}_{m}_{seed}"] = L
        out[f"L_{dim}_{m}_{seed}_sha"] = _d(L)
        out[f"Lcodes_{dim}_{m}_{seed}"] = codes
        out[f"Lscales_{dim}_{m}_{seed}"] = scales
        qr = np.random.default_rng(dim + m).standard_normal((6, dim)).astype(np.float32)
        qr[5] = 0.0
        out[f"qrot_{dim}_{m}_{seed}"] = qr
        out[f"qland_{dim}_{m}_{seed}"] = R.project_q4(qr, L, (m + 7) // 8 * 8)
    out["L_1536_1024_0_sha"] = _d(R.ortho_landmarks(1536, 1024, 0))
    np.savez_compressed(OUT / "landmark.npz", **out)


def main():
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"], check=True, stdout=subprocess.DEVNULL)
    R = Oracle("reference")
    only = sys.argv[1] if len(sys.argv) > 1 else None
    if only in (None, "score_into"):
        score_into_fixture(R)
    if only in (None, "rescore"):
        rescore_fixture(R)
    if only in (None, "score_prepared"):
        score_prepared_fixture(R)
    if only in (None, "landmark"):
        landmark_fixture(R)
    if only is not None:
        print("wrote", only)
        return

    rot = {}
    for d in ROT_DIMS:
        perm, signs, isb, blk = R.rotation_tables(d)
        rot[f"perm_{d}"] = perm
        rot[f"signs_{d}"] = signs
        rot[f"isb_{d}"] = np.array([isb], np.float32)
        rot[f"block_{d}"] = np.array([blk], np.int64)
    r = np.random.default_rng(11)
    for d in (128, 200, 1536):
        row = (r.standard_normal(d) * 3.0).astype(np.float32)
        rot[f"rowin_{d}"] = row
        rot[f"norm_{d}"] = np.array([R.row_norm(row)], np.float32)
        rot[f"rotated_{d}"] = R.rotate_scaled(row, np.float32(1.0) / np.float32(R.row_norm(row)))
    np.savez_compressed(OUT / "rotation.npz", **rot)

    cb = {}
    for bits in (2, 3, 4):
        for d in CB_DIMS:
            b, c = R.codebook(bits, d)
            cb[f"bnd_{bits}_{d}"] = b
            cb[f"cen_{bits}_{d}"] = c
    for a in (3.5, 63.5, 383.5, 767.5):
        cb[f"betacdf_{a}"] = np.array([R.beta_cdf_aa(a, t) for t in (0.3, 0.5, 0.52, 0.6, 0.9)])
    np.savez_compressed(OUT / "codebook.npz", **cb)

    manifest = {"cases": [], "ks": KS}
    for idx, (bits, dim, n, tq) in enumerate(CASES):
        x, q, sample = case_data(bits, dim, n, idx)
        f = {"x_sha": _d(x), "q_sha": _d(q)}
        shift = scale = None
        if tq:
            shift, scale = R.fit_calibration(sample, bits)
            f.update(sample_sha=_d(sample), shift=shift, scale=scale)
        packed, alpha, rotated = R.encode(x, bits, shift, scale, want_rotated=True)
        flat = R.codes_flat(packed, bits, dim)
        perm0 = R.repack_perm0(packed, bits, dim)
        seq = R.repack_seq(packed, bits, dim)
        prep = R.prepare_queries(q, bits, shift, scale)
        f.update(packed=packed, alpha=alpha, flat=flat, perm0_sha=_d(perm0), seq_sha=_d(seq),
                 rotated_sha=_d(rotated))
        for key, val in prep.items():
            if key == "q_rot":
                f["prep_q_rot_sha"] = _d(val)
            else:
                f["prep_" + key] = val
        for k in KS:
            s, i = R.search_flat(q, k, bits, perm0, alpha, shift, scale)
            f[f"scores_k{k}"] = s
            f[f"ids_k{k}"] = i
        # multi-bucket plumbing: scoring with an id_map (no TQ+)
        if not tq:
            id_map = (np.arange(n, dtype=np.uint64)[::-1] * 3 + 5).copy()
            s, i = _idmap_search(R, q, 10, bits, perm0, alpha, id_map)
            f.update(idmap=id_map, idmap_scores=s, idmap_ids=i)
        np.savez_compressed(OUT / f"case_{idx}.npz", **f)
        manifest["cases"].append({"idx": idx, "bits": bits, "dim": dim, "n": n, "tq": tq})

    # merge_search_results on two result sets with overlapping ids and short rows
    rm = np.random.default_rng(21)
    ds = np.sort(rm.standard_normal((5, 8)).astype(np.float32), axis=1)[:, ::-1].copy()
    di = rm.integers(0, 20, (5, 8)).astype(np.uint64)
    ss = np.sort(rm.standard_normal((5, 6)).astype(np.float32), axis=1)[:, ::-1].copy()
    si = rm.integers(0, 20, (5, 6)).astype(np.uint64)
    ms, mi = R.merge_results(ds, di, ss, si)
    np.savez_compressed(OUT / "merge.npz", ds=ds, di=di, ss=ss, si=si, ms=ms, mi=mi)

    # full-size digests of the 100k x 128, 4-bit workload (1000 queries, k = 10)
    x = synthetic.flat100k_db()
    q = synthetic.flat100k_queries()
    packed, alpha = R.encode(x, 4)
    perm0 = R.repack_perm0(packed, 4, 128)
    prep = R.prepare_queries(q, 4)
    s, i = R.search_flat(q, 10, 4, perm0, alpha)
    manifest["flat100k"] = {
        "x": sha(x), "q": sha(q), "packed": sha(packed), "alpha": sha(alpha), "perm0": sha(perm0),
        "luts": sha(prep["luts"]), "lut_scale": sha(prep["lut_scale"]), "lut_bias": sha(prep["lut_bias"]),
        "scores_k10": sha(s), "ids_k10": sha(i),
    }
    manifest["generator"] = {
        "reference_build": "g++ -O3 -DNDEBUG -std=c++20 -mavx2 -mfma -ffp-contract=fast -fopenmp (oracle/Makefile)",
        "env": "TURBOVEC_NO_VNNI=1 (AVX2 PERM0 uint8-LUT scan)",
        "numpy": np.__version__,
    }
    (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1))
    print("wrote", sorted(p.name for p in OUT.glob("*.npz")))


def _idmap_search(R, q, k, bits, perm0, alpha, id_map):
    import ctypes

    q = np.ascontiguousarray(q, np.float32)
    nq, dim = q.shape
    n = alpha.size
    kk = min(k, n)
    scores = np.zeros(nq * kk, np.float32)
    ids = np.zeros(nq * kk, np.uint64)
    ko = ctypes.c_size_t()
    P, S = ctypes.c_void_p, ctypes.c_size_t
    rc = R.lib.vcref_search_flat_idmap(q.ctypes.data_as(P), S(nq), S(dim), S(k), S(bits),
                                       perm0.ctypes.data_as(P), S((n + 191) // 192),
                                       alpha.ctypes.data_as(P), S(n), id_map.ctypes.data_as(P),
                                       scores.ctypes.data_as(P), ids.ctypes.data_as(P), ctypes.byref(ko))
    assert rc == 0
    return scores.reshape(nq, ko.value), ids.reshape(nq, ko.value)


if __name__ == "__main__":
    main()
