"""Pins the CPU oracle (oracle/liboracle.so) before anything is checked against it.

1. Golden vectors written from the reference's own compiled code
   (tests/golden/, make_golden.py): every oracle output is bit-identical.
2. The reference build itself (oracle/_ref, when present): random shapes.
3. What the mathematics fixes, independent of the reference: the ChaCha20
   block function against RFC 7539 §2.3.2, orthogonality / norm preservation
   of the rotation, codebook symmetry / midpoints / large-dim Gaussian limit,
   order statistics by brute-force sort, LUT window sums by brute force, and
   top-k against a full sort.
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import load_golden
from util import bits_equal, digest, scores_equal, sha


# ---------------------------------------------------------------- golden vectors
def test_rotation_tables_golden(oracle):
    g = load_golden("rotation.npz")
    for key in g:
        if not key.startswith("perm_"):
            continue
        d = int(key.split("_")[1])
        perm, signs, isb, blk = oracle.rotation_tables(d)
        assert np.array_equal(perm, g[f"perm_{d}"]), d
        assert bits_equal(signs, g[f"signs_{d}"]), d
        assert bits_equal(np.float32([isb]), g[f"isb_{d}"]) and blk == int(g[f"block_{d}"][0]), d
    for d in (128, 200, 1536):
        row = g[f"rowin_{d}"]
        nrm = oracle.row_norm(row)
        assert bits_equal(np.float32([nrm]), g[f"norm_{d}"]), d
        assert bits_equal(oracle.rotate_scaled(row, np.float32(1.0) / nrm), g[f"rotated_{d}"]), d


def test_codebook_golden(oracle):
    g = load_golden("codebook.npz")
    for key in g:
        if key.startswith("cen_"):
            _, bits, d = key.split("_")
            b, c = oracle.codebook(int(bits), int(d))
            assert bits_equal(c, g[key]) and bits_equal(b, g[f"bnd_{bits}_{d}"]), key
        if key.startswith("betacdf_"):
            a = float(key.split("_")[1])
            vals = np.array([oracle.beta_cdf_aa(a, t) for t in (0.3, 0.5, 0.52, 0.6, 0.9)])
            assert bits_equal(vals, g[key]), key


def test_encode_pack_prepare_search_golden(oracle, golden_cases):
    for c, g, x, q, sample in golden_cases:
        bits, dim = c["bits"], c["dim"]
        assert np.array_equal(digest(x), g["x_sha"]) and np.array_equal(digest(q), g["q_sha"]), c
        sh = sc = None
        if c["tq"]:
            assert np.array_equal(digest(sample), g["sample_sha"])
            sh, sc = oracle.fit_calibration(sample, bits)
            assert bits_equal(sh, g["shift"]) and bits_equal(sc, g["scale"]), c
        packed, alpha, rot = oracle.encode(x, bits, sh, sc, want_rotated=True)
        assert np.array_equal(packed, g["packed"]), c
        assert bits_equal(alpha, g["alpha"]), c
        assert np.array_equal(digest(rot), g["rotated_sha"]), c
        flat = oracle.codes_flat(packed, bits, dim)
        assert np.array_equal(flat, g["flat"]), c
        perm0 = oracle.repack_perm0(packed, bits, dim)
        assert np.array_equal(digest(perm0), g["perm0_sha"]), c
        assert np.array_equal(digest(oracle.repack_seq(packed, bits, dim)), g["seq_sha"]), c
        assert np.array_equal(oracle.unpack_perm0(perm0, c["n"], flat.shape[1]), flat), c
        prep = oracle.prepare_queries(q, bits, sh, sc)
        assert np.array_equal(digest(prep["q_rot"]), g["prep_q_rot_sha"]), c
        for key in ("luts", "lut_scale", "lut_bias", "bias_corr"):
            assert bits_equal(prep[key], g["prep_" + key]), (c, key)
        for k in (1, 10, 64, 1000):
            s, i = oracle.search_flat(q, k, bits, perm0, alpha, sh, sc)
            assert np.array_equal(i, g[f"ids_k{k}"]) and bits_equal(s, g[f"scores_k{k}"]), (c, k)
        if "idmap" in g:
            s, i = oracle.scan_topk(prep["luts"], prep["lut_scale"], prep["lut_bias"] + prep["bias_corr"], flat,
                                    alpha, 10, id_map=g["idmap"])
            assert np.array_equal(i, g["idmap_ids"]) and bits_equal(s, g["idmap_scores"]), c


def test_merge_golden(oracle):
    m = load_golden("merge.npz")
    s, i = oracle.merge_results(m["ds"], m["di"], m["ss"], m["si"])
    assert np.array_equal(i, m["mi"]) and bits_equal(s, m["ms"])


def test_flat100k_digests(oracle, manifest):
    from vectorcache_amd import synthetic

    want = manifest["flat100k"]
    x = synthetic.flat100k_db()
    q = synthetic.flat100k_queries()
    assert sha(x) == want["x"] and sha(q) == want["q"]
    packed, alpha = oracle.encode(x, 4)
    assert sha(packed) == want["packed"] and sha(alpha) == want["alpha"]
    perm0 = oracle.repack_perm0(packed, 4, 128)
    assert sha(perm0) == want["perm0"]
    prep = oracle.prepare_queries(q[:1000], 4)
    assert sha(prep["luts"]) == want["luts"]
    s, i = oracle.search_flat(q, 10, 4, perm0, alpha)
    assert sha(i) == want["ids_k10"] and sha(s) == want["scores_k10"]


# ----------------------------------------------------------- the reference build
@pytest.mark.parametrize("bits,dim,n,nq,tq", [(4, 64, 500, 5, False), (2, 256, 900, 7, True),
                                              (3, 1024, 400, 4, True), (4, 1536, 700, 3, True)])
def test_oracle_vs_reference_build(oracle, ref, bits, dim, n, nq, tq):
    from vectorcache_amd import synthetic

    x = synthetic.mixture(n, dim, 24, 0.3, dim + bits, 31 + n)
    q = synthetic.mixture(nq, dim, 24, 0.3, dim + bits, 32 + n)
    sh = sc = None
    if tq:
        sample = synthetic.mixture(300, dim, 24, 0.3, dim + bits, 33)
        sh, sc = ref.fit_calibration(sample, bits)
        sh2, sc2 = oracle.fit_calibration(sample, bits)
        assert bits_equal(sh, sh2) and bits_equal(sc, sc2)
    pr, ar = ref.encode(x, bits, sh, sc)
    po, ao = oracle.encode(x, bits, sh, sc)
    assert np.array_equal(pr, po) and bits_equal(ar, ao)
    blk = ref.repack_perm0(pr, bits, dim)
    assert np.array_equal(blk, oracle.repack_perm0(po, bits, dim))
    for k in (3, 50, n + 5):
        s1, i1 = ref.search_flat(q, k, bits, blk, ar, sh, sc)
        s2, i2 = oracle.search_flat(q, k, bits, blk, ar, sh, sc)
        assert np.array_equal(i1, i2) and bits_equal(s1, s2), k


def test_codebooks_all_dims_vs_reference_build(oracle, ref):
    for bits in (2, 3, 4):
        for d in list(range(8, 2049, 8)) + [3072, 4096, 8192, 16384]:
            a, b = oracle.codebook(bits, d), ref.codebook(bits, d)
            assert bits_equal(a[1], b[1]) and bits_equal(a[0], b[0]), (bits, d)


# ------------------------------------------------- what the mathematics fixes
def _block(oracle, words, rounds):
    fn = oracle.lib.oc_chacha_block
    inp = (ctypes.c_uint32 * 16)(*words)
    out = (ctypes.c_uint32 * 16)()
    fn(inp, out, ctypes.c_int(rounds))
    return list(out)


def test_chacha20_block_rfc7539(oracle):
    """RFC 7539 §2.3.2: key 00..1f, counter 1, nonce 00000009 0000004a 00000000."""
    key = bytes(range(32))
    words = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    words += [int.from_bytes(key[4 * i:4 * i + 4], "little") for i in range(8)]
    nonce = bytes.fromhex("000000090000004a00000000")
    words += [1] + [int.from_bytes(nonce[4 * i:4 * i + 4], "little") for i in range(3)]
    expect = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204, 0x4E6CD4C3,
              0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]
    assert _block(oracle, words, 20) == expect


@pytest.mark.parametrize("dim", [8, 200, 256, 768, 1024, 1536])
def test_rotation_is_orthonormal(oracle, dim):
    perm, signs, isb, blk = oracle.rotation_tables(dim)
    assert sorted(perm.tolist()) == list(range(dim)) and set(np.unique(signs)) <= {-1.0, 1.0}
    assert blk == dim & -dim and blk & (blk - 1) == 0
    m = np.stack([oracle.rotate_scaled(e, 1.0) for e in np.eye(dim, dtype=np.float32)]).astype(np.float64)
    assert np.abs(m @ m.T - np.eye(dim)).max() < 1e-5
    r = np.random.default_rng(dim)
    v = r.standard_normal(dim).astype(np.float32)
    assert abs(np.linalg.norm(oracle.rotate_scaled(v, 1.0)) - np.linalg.norm(v)) < 1e-5 * np.linalg.norm(v)
    # apply with a scale == apply to the scaled row (the reference's test_rotation.cpp:29-51)
    inv = np.float32(1.0) / oracle.row_norm(v)
    assert bits_equal(oracle.rotate_scaled(v, inv), oracle.rotate_scaled((v * inv).astype(np.float32), 1.0))


# Lloyd-Max optimal output levels of N(0, 1) (J. Max, 1960), positive half
_MAX_LEVELS = {2: [0.4528, 1.5104], 3: [0.2451, 0.7560, 1.3439, 2.1520],
               4: [0.1284, 0.3881, 0.6568, 0.9424, 1.2562, 1.6181, 2.0690, 2.7326]}


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_codebook_properties(oracle, bits):
    for d in (8, 128, 1536, 16384):
        b, c = oracle.codebook(bits, d)
        assert np.all(np.diff(c) > 0)
        assert bits_equal(b, ((c[:-1] + c[1:]) * np.float32(0.5)).astype(np.float32))
        assert np.abs(c + c[::-1]).max() <= 2e-7 * np.abs(c).max()
    # Beta((d-1)/2,(d-1)/2) on [-1,1] -> N(0, 1/d): sqrt(d) * levels -> Max's levels
    _, c = oracle.codebook(bits, 16384)
    assert np.allclose(c[len(c) // 2:] * math.sqrt(16384), _MAX_LEVELS[bits], atol=3e-3)


def test_calibration_order_statistics_brute_force(oracle):
    from vectorcache_amd import synthetic

    bits, dim, n = 4, 256, 1001
    x = synthetic.mixture(n, dim, 8, 0.3, 5, 6)
    _, _, rot = oracle.encode(x, bits, want_rotated=True)
    _, cen = oracle.codebook(bits, dim)
    c_outer = np.float32(np.abs(cen).max())
    a = (dim - 1) / 2.0
    p_hi = oracle.beta_cdf_aa(a, (float(c_outer) + 1.0) / 2.0)
    lo_idx = int(n * (1.0 - p_hi))
    hi_idx = max(min(int(n * p_hi), n - 1), lo_idx + 1)
    srt = np.sort(rot, axis=0)
    q_lo, q_hi = srt[lo_idx], srt[hi_idx]
    span = (q_hi - q_lo).astype(np.float32)
    scale = (np.float32(2 * c_outer) / span).astype(np.float32)
    shift = ((-c_outer) / scale - q_lo).astype(np.float32)
    sh, sc = oracle.fit_calibration(x, bits)
    assert bits_equal(sc, scale) and bits_equal(sh, shift)


def test_window_sums_brute_force(oracle):
    r = np.random.default_rng(9)
    G = 600  # three windows, the last one partial
    lut = r.integers(0, 128, G * 32).astype(np.uint8)
    flat = r.integers(0, 256, (50, G)).astype(np.uint8)
    got = oracle.window_sums(lut, flat)
    hi = lut.reshape(G, 32)[np.arange(G), flat >> 4].astype(np.int64)
    lo = lut.reshape(G, 32)[np.arange(G), 16 + (flat & 15)].astype(np.int64)
    want = np.stack([(hi + lo)[:, w * 256:(w + 1) * 256].sum(1) for w in range(3)], 1)
    assert np.array_equal(got, want)


def test_topk_is_full_sort(oracle):
    """The oracle's top-k equals a full sort by (score desc, id asc) of its own
    per-vector scores (k = n returns every score)."""
    r = np.random.default_rng(4)
    G, n = 64, 300
    lut = r.integers(0, 8, (2, G * 32)).astype(np.uint8)    # small values -> many ties
    flat = r.integers(0, 256, (n, G)).astype(np.uint8)
    alpha = np.ones(n, np.float32)
    scale = np.float32([0.5, 0.25])
    bias = np.float32([0.0, -1.0])
    s_all, i_all = oracle.scan_topk(lut, scale, bias, flat, alpha, n)
    for qi in range(2):
        order = np.lexsort((i_all[qi], -s_all[qi]))
        assert np.array_equal(order, np.arange(n))           # already sorted
        assert sorted(i_all[qi].tolist()) == list(range(n))  # a permutation
        s10, i10 = oracle.scan_topk(lut[qi:qi + 1], scale[qi:qi + 1], bias[qi:qi + 1], flat, alpha, 10)
        assert np.array_equal(i10[0], i_all[qi][:10])
        sums = oracle.window_sums(lut[qi], flat)[:, 0]
        want = (scale[qi] * sums.astype(np.float64) + bias[qi]).astype(np.float32)  # exact here
        assert bits_equal(np.sort(want)[::-1], s_all[qi])


# ------------------------------------------------- score_prepared_into (heaps)
def test_score_into_golden(oracle):
    """Chained score_prepared_into calls (several buckets, id maps, query
    subsets, offsets, a seeded threshold; rescan heap k < 128 and binary heap
    k >= 128): every heap array after every call equals the reference's."""
    from util import empty_heaps, heaps_equal, score_into_scenarios

    g = load_golden("score_into.npz")
    n_checked = 0
    for name, bits, dim, k, q, shift, scale, calls, want in score_into_scenarios(g):
        heaps = empty_heaps(q.shape[0], k, {4: 0.05})
        for ci, (perm0, alpha, id_map, qidx, off) in enumerate(calls):
            heaps = oracle.score_into(q, k, bits, perm0, alpha, qidx, heaps, id_map=id_map, shift=shift,
                                      scale=scale, offsets=off)
            assert heaps_equal(heaps, want[ci]), (name, k, ci)
            n_checked += 1
        s, i = oracle.heaps_to_results(heaps[0], heaps[1], heaps[2])
        assert bits_equal(s, want["final"][0]) and np.array_equal(i, want["final"][1]), (name, k)
    assert n_checked >= 10


@pytest.mark.parametrize("k", [5, 40, 200])
def test_score_into_single_bucket_is_full_sort(oracle, k):
    """One bucket into empty heaps (threshold -inf), ids increasing with the
    scan order: strict admission in scan order keeps, among equal scores, the
    earlier (lower) id, so the heaps hold the exact top-k of a full sort by
    (score desc, id asc) -- an independent check of the heap semantics."""
    from vectorcache_amd import synthetic

    from util import empty_heaps

    n, dim, bits = 700, 64, 4
    x = synthetic.mixture(n, dim, 16, 0.3, 91, 92)
    x[100:140] = x[99]                       # ties
    q = synthetic.mixture(4, dim, 16, 0.3, 91, 93)
    packed, alpha = oracle.encode(x, bits)
    perm0 = oracle.repack_perm0(packed, bits, dim)
    id_map = (np.arange(n, dtype=np.uint64) * 5 + 2)
    heaps = oracle.score_into(q, k, bits, perm0, alpha, np.arange(4, dtype=np.uint64), empty_heaps(4, k),
                              id_map=id_map)
    s, i = oracle.heaps_to_results(heaps[0], heaps[1], heaps[2])
    prep = oracle.prepare_queries(q, bits)
    flat = oracle.codes_flat(packed, bits, dim)
    s_all, i_all = oracle.scan_topk(prep["luts"], prep["lut_scale"], prep["lut_bias"] + prep["bias_corr"], flat,
                                    alpha, n)
    for qi in range(4):
        order = np.lexsort((i_all[qi], -s_all[qi]))[:k]
        assert np.array_equal(i[qi], id_map[i_all[qi][order]]), (k, qi)
        assert scores_equal(s[qi], s_all[qi][order]), (k, qi)


def test_score_into_vs_reference_build(oracle, ref):
    """Random chained calls, the oracle against the reference build."""
    from vectorcache_amd import synthetic

    from util import empty_heaps, heaps_equal

    r = np.random.default_rng(77)
    for bits, dim, k in ((4, 128, 9), (3, 96, 130), (2, 64, 33)):
        n = 420
        x = synthetic.mixture(n, dim, 8, 0.2, dim, bits)
        x[r.integers(0, n, 30)] = x[5]
        q = synthetic.mixture(5, dim, 8, 0.2, dim, bits + 50)
        packed, alpha = ref.encode(x, bits)
        ho, hr = empty_heaps(5, k, {2: -0.25}), empty_heaps(5, k, {2: -0.25})
        for ci, (a, b) in enumerate(((0, 160), (160, 161), (161, 420))):
            perm0 = ref.repack_perm0(packed[a:b], bits, dim)
            qidx = r.permutation(5)[: 2 + ci].astype(np.uint64)
            off = (r.standard_normal(5) * (ci % 2)).astype(np.float32)
            idm = r.permutation(10 * n)[: b - a].astype(np.uint64)
            hr = ref.score_into(q, k, bits, perm0, alpha[a:b], qidx, hr, id_map=idm, offsets=off)
            ho = oracle.score_into(q, k, bits, perm0, alpha[a:b], qidx, ho, id_map=idm, offsets=off)
            assert heaps_equal(hr, ho), (bits, dim, k, ci)


# ------------------------------------------------------------ float re-rank
def test_float_rescore_golden(oracle):
    """float_rescore_topk: every dim's rows (all entries, and per-row counts)
    equal the reference's, bit for bit; ids past the corpus score -inf."""
    from cases import RESCORE_DIMS, rescore_inputs

    g = load_golden("float_rescore.npz")
    for dim in RESCORE_DIMS:
        corpus, q, scores, ids, counts = rescore_inputs(dim)
        for tag, cnt in (("all", None), ("counts", counts)):
            s, i = oracle.float_rescore(scores, ids, q, corpus, cnt)
            assert bits_equal(s, g[f"d{dim}_{tag}_scores"]) and np.array_equal(i, g[f"d{dim}_{tag}_ids"]), (dim, tag)


def test_float_rescore_is_the_inner_product(oracle):
    """Each re-scored entry is the float inner product to within the error of
    a float sum of dim products, and rows come out sorted (score desc, id asc)."""
    r = np.random.default_rng(8)
    dim = 1536
    corpus = r.standard_normal((200, dim)).astype(np.float32)
    q = r.standard_normal((4, dim)).astype(np.float32)
    ids = r.integers(0, 200, (4, 20)).astype(np.uint64)
    s, i = oracle.float_rescore(np.zeros((4, 20), np.float32), ids, q, corpus)
    exact = np.einsum("qd,qkd->qk", q.astype(np.float64), corpus[i.astype(np.int64)].astype(np.float64))
    bound = dim * 2.0 ** -23 * np.einsum("qd,qkd->qk", np.abs(q), np.abs(corpus[i.astype(np.int64)]))
    assert np.all(np.abs(s - exact) <= bound)
    for row_s, row_i in zip(s, i):
        order = np.lexsort((row_i, -row_s))
        assert np.array_equal(order, np.arange(20))


def test_gpu32_layout_known_answer(oracle):
    """The reference CUDA index's code words (src/gpu/kernels.cu:111-133),
    written out by hand for 33 vectors of 5 byte groups: 2 tiles of 32 lanes,
    2 words per lane (groups 0-3, then group 4 alone), little-endian bytes,
    zero past the last group and the last vector."""
    n, G = 33, 5
    flat = (np.arange(n * G, dtype=np.uint32) % 251 + 1).astype(np.uint8).reshape(n, G)
    w = oracle.flat_to_gpu32(flat)
    assert w.size == 128
    b = flat.astype(np.uint32)
    # vector 0: tile 0, lane 0 -> words 0 (groups 0-3) and 32 (group 4)
    assert w[0] == b[0, 0] | b[0, 1] << 8 | b[0, 2] << 16 | b[0, 3] << 24
    assert w[32] == b[0, 4]
    # vector 1: lane 1 -> words 1 and 33
    assert w[1] == b[1, 0] | b[1, 1] << 8 | b[1, 2] << 16 | b[1, 3] << 24
    assert w[33] == b[1, 4]
    # vector 31: lane 31 -> words 31 and 63
    assert w[31] == b[31, 0] | b[31, 1] << 8 | b[31, 2] << 16 | b[31, 3] << 24
    assert w[63] == b[31, 4]
    # vector 32: tile 1, lane 0 -> words 64 and 96
    assert w[64] == b[32, 0] | b[32, 1] << 8 | b[32, 2] << 16 | b[32, 3] << 24
    assert w[96] == b[32, 4]
    # lanes 1-31 of tile 1 hold no vector
    assert not w[65:96].any() and not w[97:128].any()


def test_gpu32_round_trip(oracle):
    r = np.random.default_rng(17)
    for n, G in [(1, 1), (31, 3), (32, 4), (1000, 24), (257, 384)]:
        flat = r.integers(0, 256, (n, G), dtype=np.uint8)
        w = oracle.flat_to_gpu32(flat)
        assert w.size == (n + 31) // 32 * ((G + 3) // 4) * 32
        assert np.array_equal(oracle.gpu32_to_flat(w, n, G), flat)
        # a bijection between the valid bytes: every flat byte appears once
        assert int(np.count_nonzero(w.view(np.uint8))) == int(np.count_nonzero(flat))


def _score_prepared_codes(oracle, bits, dim, n, nq):
    from cases import score_prepared_inputs

    x, q, id_map, qi, rn, rs = score_prepared_inputs(bits, dim, n, nq)
    packed, alpha = oracle.encode(x, bits)
    return q, oracle.repack_perm0(packed, bits, dim), alpha, (id_map, qi, rn, rs)


def test_score_prepared_golden(oracle):
    """score_prepared with id_map, query subsets (with a repeat) and the
    residual post-correction: bit-identical to the reference build's outputs."""
    from cases import SCORE_PREPARED_CASES, SCORE_PREPARED_VARIANTS

    g = load_golden("score_prepared.npz")
    for ci, (bits, dim, n, nq, k) in enumerate(SCORE_PREPARED_CASES):
        q, perm0, alpha, extra = _score_prepared_codes(oracle, bits, dim, n, nq)
        for tag, kw in SCORE_PREPARED_VARIANTS(*extra):
            s, i = oracle.score_prepared(q, k, bits, perm0, alpha, **kw)
            assert np.array_equal(i, g[f"c{ci}_{tag}_ids"]), (ci, tag)
            assert bits_equal(s, g[f"c{ci}_{tag}_scores"]), (ci, tag)


def test_score_prepared_residual_is_a_rerank(oracle):
    """What the mathematics fixes: the correction keeps the LUT top-k
    membership, adds rs * ||r|| per kept vector (one rounding) and re-sorts."""
    q, perm0, alpha, (id_map, qi, rn, rs) = _score_prepared_codes(oracle, 4, 128, 3000, 5)
    s0, i0 = oracle.score_prepared(q, 50, 4, perm0, alpha)
    s1, i1 = oracle.score_prepared(q, 50, 4, perm0, alpha, residual_norms=rn, residual_scales=rs)
    for r in range(q.shape[0]):
        assert set(i0[r].tolist()) == set(i1[r].tolist())
        exact = {int(v): Fraction(float(s)) + Fraction(float(rs[r])) * Fraction(float(rn[v]))
                 for s, v in zip(s0[r], i0[r])}
        for s, v in zip(s1[r], i1[r]):
            # s is the float32 nearest the exact s0 + rs * ||r|| (one rounding)
            e = exact[int(v)]
            lo, hi = np.nextafter(s, np.float32(-np.inf)), np.nextafter(s, np.float32(np.inf))
            d = abs(Fraction(float(s)) - e)
            assert d <= abs(Fraction(float(lo)) - e) and d <= abs(Fraction(float(hi)) - e), (r, int(v))
        assert np.all(np.diff(s1[r]) <= 0)


def test_score_prepared_vs_reference_build(oracle, ref):
    for bits, dim, n, nq, k in [(4, 256, 2500, 6, 33), (2, 128, 900, 3, 900)]:
        q, perm0, alpha, (id_map, qi, rn, rs) = _score_prepared_codes(oracle, bits, dim, n, nq)
        for kw in ({}, dict(id_map=id_map, query_indices=qi, residual_norms=rn, residual_scales=rs)):
            a = oracle.score_prepared(q, k, bits, perm0, alpha, **kw)
            b = ref.score_prepared(q, k, bits, perm0, alpha, **kw)
            assert np.array_equal(a[1], b[1]) and bits_equal(a[0], b[0]), (bits, dim, list(kw))
