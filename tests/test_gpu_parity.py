"""GPU parity: the HIP path (through the C ABI) against the reference's golden
vectors and against the CPU oracle on the same seeded inputs.

Bar: bit-exact codes, PERM0 blocks, alpha, TQ+ shift/scale, LUT bytes/scale/
bias, scores and top-k ids (ties by lower id). Scores compare with -0.0 ==
+0.0 (the library reports -0.0 as +0.0), every other bit must match.
"""
import numpy as np
import pytest

from conftest import load_golden
from util import bits_equal, scores_equal, sha

pytestmark = pytest.mark.gpu


def _index(vc, c, x, g, sample):
    ix = vc.TurboQuantIndex(c["dim"], c["bits"])
    if c["tq"]:
        ix.calibrate(sample)
    ix.add(x)
    ix.prepare()
    return ix


def test_version(vc):
    assert "gfx950" in vc.version()


def test_rotation_matches_golden(vc, golden_cases):
    from conftest import load_golden

    g = load_golden("rotation.npz")
    for d in (128, 200, 1536):
        rot = vc.rotate(g[f"rowin_{d}"][None, :])
        assert bits_equal(rot[0], g[f"rotated_{d}"]), d
    for c, gc, x, q, sample in golden_cases:
        assert sha(vc.rotate(x)) == bytes(gc["rotated_sha"]).hex(), c


def test_encode_matches_golden(vc, golden_cases):
    for c, g, x, q, sample in golden_cases:
        sh, sc = (g["shift"], g["scale"]) if c["tq"] else (None, None)
        packed, alpha = vc.encode(x, c["bits"], sh, sc)
        assert np.array_equal(packed, g["packed"]), c
        assert bits_equal(alpha, g["alpha"]), c


def test_fit_calibration_matches_golden(vc, golden_cases):
    for c, g, x, q, sample in golden_cases:
        if not c["tq"]:
            continue
        sh, sc = vc.fit_calibration(sample, c["bits"])
        assert bits_equal(sh, g["shift"]), c
        assert bits_equal(sc, g["scale"]), c


def test_index_exports_and_search_match_golden(vc, golden_cases):
    for c, g, x, q, sample in golden_cases:
        ix = _index(vc, c, x, g, sample)
        assert ix.size() == c["n"]
        if c["tq"]:
            sh, sc = ix.calibration()
            assert bits_equal(sh, g["shift"]) and bits_equal(sc, g["scale"]), c
        assert np.array_equal(ix.export_packed(), g["packed"]), c
        assert np.array_equal(ix.export_flat(), g["flat"]), c
        assert sha(ix.export_perm0()) == bytes(g["perm0_sha"]).hex(), c
        assert sha(ix.export_seq()) == bytes(g["seq_sha"]).hex(), c
        assert bits_equal(ix.export_alpha(), g["alpha"]), c
        prep = ix.prepare_queries(q)
        assert np.array_equal(prep["luts"], g["prep_luts"]), c
        assert bits_equal(prep["lut_scale"], g["prep_lut_scale"]), c
        assert bits_equal(prep["lut_bias"], g["prep_lut_bias"]), c
        assert bits_equal(prep["bias_corr"], g["prep_bias_corr"]), c
        for k in (1, 10, 64, 1000):
            r = ix.search(q, k)
            assert np.array_equal(r.ids, g[f"ids_k{k}"]), (c, k)
            assert scores_equal(r.scores, g[f"scores_k{k}"]), (c, k)


def test_gpu32_layout_export_and_import(vc, oracle, golden_cases):
    """The reference CUDA index's code words: export equals the oracle's
    layout of the reference's flat bytes; imported back as PERM0 blocks they
    equal the PERM0 export and search like the index itself."""
    for c, g, x, q, sample in golden_cases:
        ix = _index(vc, c, x, g, sample)
        w = ix.export_gpu32()
        assert np.array_equal(w, oracle.flat_to_gpu32(g["flat"])), c
        p = vc.gpu32_to_perm0(w, c["n"], c["dim"], c["bits"])
        assert np.array_equal(p, ix.export_perm0()), c
        sh, sc = (g["shift"], g["scale"]) if c["tq"] else (None, None)
        r = vc.search_perm0(q, 10, c["bits"], p, ix.export_alpha(), shift=sh, scale=sc)
        assert np.array_equal(r.ids, g["ids_k10"]), c
        assert scores_equal(r.scores, g["scores_k10"]), c


def test_score_prepared_arguments_match_golden(vc, oracle):
    """score_prepared with id_map, query subsets (with a repeat) and the
    residual post-correction, against the reference build's golden outputs;
    then a larger random case against the oracle."""
    from cases import SCORE_PREPARED_CASES, SCORE_PREPARED_VARIANTS, score_prepared_inputs

    g = load_golden("score_prepared.npz")
    for ci, (bits, dim, n, nq, k) in enumerate(SCORE_PREPARED_CASES):
        x, q, *extra = score_prepared_inputs(bits, dim, n, nq)
        packed, alpha = oracle.encode(x, bits)
        perm0 = oracle.repack_perm0(packed, bits, dim)
        for tag, kw in SCORE_PREPARED_VARIANTS(*extra):
            r = vc.score_perm0(q, k, bits, perm0, alpha, **kw)
            assert np.array_equal(r.ids, g[f"c{ci}_{tag}_ids"]), (ci, tag)
            assert bits_equal(r.scores, g[f"c{ci}_{tag}_scores"]), (ci, tag)
    x, q, id_map, qi, rn, rs = score_prepared_inputs(4, 768, 20000, 9)
    packed, alpha = oracle.encode(x, 4)
    perm0 = oracle.repack_perm0(packed, 4, 768)
    qi = np.array([8, 3, 3, 0, 5], np.uint64)
    kw = dict(id_map=id_map, query_indices=qi, residual_norms=rn, residual_scales=rs)
    for k in (10, 700):
        s_o, i_o = oracle.score_prepared(q, k, 4, perm0, alpha, **kw)
        r = vc.score_perm0(q, k, 4, perm0, alpha, **kw)
        assert np.array_equal(r.ids, i_o) and bits_equal(r.scores, s_o), k
    with pytest.raises(IndexError):
        vc.score_perm0(q, 10, 4, perm0, alpha, query_indices=np.array([9], np.uint64))


def test_sequential_bias_corr_matches_golden(vc, golden_cases):
    """Bit 4: query preparation always through the reference's sequential
    bias_corr chain (the path the certified parallel sum falls back to):
    the same LUTs, scales, biases and bias_corr."""
    vc.set_debug(16)
    try:
        for c, g, x, q, sample in golden_cases:
            if not c["tq"]:
                continue
            ix = _index(vc, c, x, g, sample)
            prep = ix.prepare_queries(q)
            assert np.array_equal(prep["luts"], g["prep_luts"]), c
            assert bits_equal(prep["lut_bias"], g["prep_lut_bias"]), c
            assert bits_equal(prep["bias_corr"], g["prep_bias_corr"]), c
    finally:
        vc.set_debug(0)


@pytest.mark.parametrize("flags", [1, 4, 5, 64, 65])
def test_exact_alpha_path_matches_golden(vc, golden_cases, flags):
    """Bit 0: every vector through the reference's sequential alpha chains
    (the path the certified fast alpha falls back to); bit 2: the 8-rows-per-
    wavefront tile encode instead of the streaming encode; bit 6: the two-pass
    streaming encode (norm pass, row pass) instead of the one-pass kernel.
    Same bits."""
    vc.set_debug(flags)
    try:
        for c, g, x, q, sample in golden_cases:
            ix = _index(vc, c, x, g, sample)
            assert bits_equal(ix.export_alpha(), g["alpha"]), c
            assert np.array_equal(ix.export_packed(), g["packed"]), c
    finally:
        vc.set_debug(0)


@pytest.mark.parametrize("kernel,tile", [("pair", 1), ("pair", 2), ("pair", 4), ("pair", 8),
                                         ("nibble", 1), ("nibble", 4), ("nibble", 8)])
def test_scan_tiles_agree_with_golden(vc, golden_cases, kernel, tile):
    """Every scan kernel (byte-pair tables, nibble tables) at every query tile."""
    vc.set_scan_tile(tile)
    vc.set_debug(2 if kernel == "nibble" else 0)
    try:
        for c, g, x, q, sample in golden_cases:
            ix = _index(vc, c, x, g, sample)
            assert vc.scan_tile_for(c["dim"], c["bits"], q.shape[0]) == tile
            for k in (10, 64, 1000):
                r = ix.search(q, k)
                assert np.array_equal(r.ids, g[f"ids_k{k}"]), (c, k, kernel, tile)
                assert scores_equal(r.scores, g[f"scores_k{k}"]), (c, k, kernel, tile)
    finally:
        vc.set_scan_tile(0)
        vc.set_debug(0)


def test_idmap_scoring_matches_golden(vc, golden_cases):
    for c, g, x, q, sample in golden_cases:
        if "idmap" not in g:
            continue
        # the PERM0 bytes the golden run scanned are the GPU's own export
        ix = _index(vc, c, x, g, sample)
        perm0 = ix.export_perm0()
        r = vc.search_perm0(q, 10, c["bits"], perm0, g["alpha"], id_map=g["idmap"])
        assert np.array_equal(r.ids, g["idmap_ids"]), c
        assert scores_equal(r.scores, g["idmap_scores"]), c


def test_merge_matches_golden(vc):
    from conftest import load_golden

    m = load_golden("merge.npz")
    dst = vc.SearchResults(m["ds"], m["di"])
    src = vc.SearchResults(m["ss"], m["si"])
    out = vc.merge_search_results(dst, src)
    assert np.array_equal(out.ids, m["mi"])
    assert bits_equal(out.scores, m["ms"])


def test_flat100k_full_size_digests(vc, manifest, oracle):
    """The 100k x 128 4-bit workload at full size: every output's SHA-256
    equals the reference's."""
    from vectorcache_amd import synthetic

    want = manifest["flat100k"]
    x = synthetic.flat100k_db()
    q = synthetic.flat100k_queries()
    assert sha(x) == want["x"] and sha(q) == want["q"]
    ix = vc.TurboQuantIndex(128, 4)
    ix.add(x)
    ix.prepare()
    assert sha(ix.export_packed()) == want["packed"]
    assert sha(ix.export_alpha()) == want["alpha"]
    assert sha(ix.export_perm0()) == want["perm0"]
    prep = ix.prepare_queries(q)
    assert sha(prep["luts"]) == want["luts"]
    assert sha(prep["lut_scale"]) == want["lut_scale"]
    assert sha(prep["lut_bias"]) == want["lut_bias"]
    r = ix.search(q, 10)
    assert sha(r.ids) == want["ids_k10"]
    assert sha(r.scores) == want["scores_k10"]


def _oracle_search(oracle, x, q, bits, k, shift=None, scale=None):
    packed, alpha = oracle.encode(x, bits, shift, scale)
    flat = oracle.codes_flat(packed, bits, x.shape[1])
    prep = oracle.prepare_queries(q, bits, shift, scale)
    return oracle.scan_topk(prep["luts"], prep["lut_scale"], prep["lut_bias"] + prep["bias_corr"], flat, alpha, k)


@pytest.mark.parametrize("bits,dim,n,nq,k", [(4, 768, 20000, 37, 100), (2, 1536, 7000, 9, 64),
                                             (3, 384, 5000, 70, 1), (4, 1536, 3000, 5, 2048)])
def test_random_vs_oracle(vc, oracle, bits, dim, n, nq, k):
    from vectorcache_amd import synthetic

    x = synthetic.mixture(n, dim, 64, 0.3, 11 + dim, 12 + n)
    q = synthetic.mixture(nq, dim, 64, 0.3, 11 + dim, 13 + nq)
    s_o, i_o = _oracle_search(oracle, x, q, bits, k)
    ix = vc.TurboQuantIndex(dim, bits)
    ix.add(x)
    ix.prepare()
    r = ix.search(q, k)
    assert np.array_equal(r.ids, i_o)
    assert scores_equal(r.scores, s_o)


@pytest.mark.parametrize("kernel,tile,ranges", [("nibble", 1, 8), ("nibble", 8, 2), ("pair", 1, 8),
                                                ("pair", 2, 8), ("pair", 4, 3), ("pair", 8, 2),
                                                ("pair-v", 1, 8), ("pair-v", 8, 2)])
def test_multi_step_ranges_match_oracle(vc, oracle, kernel, tile, ranges):
    """Few vector ranges (forced), so every workgroup scans many steps and runs
    many rounds of candidate admission and list merges, over several query
    tiles: the 100k x 128 4-bit workload against the oracle. "pair-v" is the
    byte-pair kernel's other vectors-per-thread variant."""
    from vectorcache_amd import synthetic

    x = synthetic.flat100k_db()
    q = synthetic.flat100k_queries()[:16]
    ix = vc.TurboQuantIndex(128, 4)
    ix.add(x)
    ix.prepare()
    vc.set_scan_tile(tile)
    vc.set_debug({"nibble": 2, "pair": 0, "pair-v": 8}[kernel] | (ranges << 16))
    try:
        for k in (10, 100):
            s_o, i_o = _oracle_search(oracle, x, q, 4, k)
            r = ix.search(q, k)
            assert np.array_equal(r.ids, i_o), (kernel, tile, ranges, k)
            assert scores_equal(r.scores, s_o), (kernel, tile, ranges, k)
    finally:
        vc.set_scan_tile(0)
        vc.set_debug(0)


def test_ties_resolve_to_lower_id(vc, oracle):
    r = np.random.default_rng(3)
    base = r.standard_normal((50, 256)).astype(np.float32)
    x = base[r.integers(0, 50, 9000)]
    q = r.standard_normal((6, 256)).astype(np.float32)
    s_o, i_o = _oracle_search(oracle, x, q, 4, 300)
    ix = vc.TurboQuantIndex(256, 4)
    ix.add(x)
    ix.prepare()
    res = ix.search(q, 300)
    assert np.array_equal(res.ids, i_o)
    assert scores_equal(res.scores, s_o)
    # within equal scores ids ascend
    for row_s, row_i in zip(res.scores, res.ids):
        same = row_s[1:] == row_s[:-1]
        assert np.all(row_i[1:][same] > row_i[:-1][same])


@pytest.mark.parametrize("nq,k", [(1, 10), (2, 64)])
def test_ties_at_the_admission_bound(vc, oracle, nq, k):
    """Heavy score ties with one or two queries (one-query tiles: wavefront
    sorts and the lower score bound they set) over many steps per workgroup:
    vectors tied with the bound are still admitted and resolve to the lower
    ids, as in the oracle."""
    r = np.random.default_rng(5)
    base = r.standard_normal((40, 128)).astype(np.float32)
    x = base[r.integers(0, 40, 60000)]
    q = r.standard_normal((nq, 128)).astype(np.float32)
    s_o, i_o = _oracle_search(oracle, x, q, 4, k)
    ix = vc.TurboQuantIndex(128, 4)
    ix.add(x)
    ix.prepare()
    vc.set_debug(2 << 16)  # two vector ranges: about 30 steps per workgroup
    try:
        res = ix.search(q, k)
    finally:
        vc.set_debug(0)
    assert np.array_equal(res.ids, i_o)
    assert scores_equal(res.scores, s_o)


def test_incremental_add_equals_bulk_and_batch_equals_single(vc):
    from vectorcache_amd import synthetic

    x = synthetic.mixture(3000, 128, 32, 0.3, 5, 6)
    q = synthetic.mixture(7, 128, 32, 0.3, 5, 7)
    a = vc.TurboQuantIndex(128, 4)
    a.add(x)
    a.prepare()
    b = vc.TurboQuantIndex(128, 4)
    b.add(x[:1000])
    b.prepare()
    b.add(x[1000:2100])
    b.add(x[2100:])
    assert not b.prepared
    b.prepare()
    ra, rb = a.search(q, 25), b.search(q, 25)
    assert np.array_equal(ra.ids, rb.ids) and bits_equal(ra.scores, rb.scores)
    for i in range(q.shape[0]):
        ri = a.search(q[i : i + 1], 25)
        assert np.array_equal(ri.ids[0], ra.ids[i]) and bits_equal(ri.scores[0], ra.scores[i])


def test_api_errors_and_edges(vc):
    from vectorcache_amd._capi import IndexStateError

    with pytest.raises(ValueError):
        vc.TurboQuantIndex(100, 4)  # not a multiple of 8
    with pytest.raises(ValueError):
        vc.TurboQuantIndex(128, 5)
    with pytest.raises(ValueError):
        vc.TurboQuantIndex(16392, 4)
    ix = vc.TurboQuantIndex(64, 4)
    x = np.random.default_rng(0).standard_normal((10, 64)).astype(np.float32)
    with pytest.raises(ValueError):
        ix.add(x.ravel()[:-3])  # size not a multiple of dim
    ix.add(x)
    with pytest.raises(IndexStateError):
        ix.search(x[:2], 3)  # before prepare
    ix.prepare()
    r = ix.search(x[:2], 50)  # k clipped to size
    assert r.k == 10 and r.ids[0, 0] == 0 and r.ids[1, 0] == 1
    with pytest.raises(IndexStateError):
        ix.calibrate(x)  # after vectors were encoded
    empty = vc.TurboQuantIndex(64, 4)
    empty.prepare()
    assert empty.search(x[:3], 5).k == 0
    with pytest.raises(ValueError):
        vc.TurboQuantIndex(64, 4).calibrate(x[:1])  # needs >= 2 rows


# ------------------------------------------------- score_prepared_into (heaps)
def test_score_into_matches_golden(vc):
    """Chained score_prepared_into calls on the GPU: every heap array (scores,
    ids, sizes, thresholds, worst positions) after every call equals the
    reference's, and heaps_to_search_results of the last state too."""
    from conftest import load_golden
    from util import empty_heaps, heaps_equal, score_into_scenarios

    g = load_golden("score_into.npz")
    checked = 0
    for name, bits, dim, k, q, shift, scale, calls, want in score_into_scenarios(g):
        heaps = empty_heaps(q.shape[0], k, {4: 0.05})
        for ci, (perm0, alpha, id_map, qidx, off) in enumerate(calls):
            vc.score_perm0_into(q, k, bits, perm0, alpha, qidx, heaps, id_map=id_map, shift=shift, scale=scale,
                                score_offsets=off)
            assert heaps_equal(heaps, want[ci]), (name, k, ci)
            checked += 1
        r = vc.heaps_to_results(heaps)
        assert bits_equal(r.scores, want["final"][0]) and np.array_equal(r.ids, want["final"][1]), (name, k)
    assert checked >= 10


@pytest.mark.parametrize("bits,dim,k,n", [(4, 256, 10, 3000), (4, 768, 300, 1500), (3, 128, 64, 2500)])
def test_score_into_matches_oracle(vc, oracle, bits, dim, k, n):
    """Larger buckets (many sub-tiles, several 256-group windows at dim 768),
    two chained calls with offsets, against the oracle."""
    from vectorcache_amd import synthetic

    from util import empty_heaps, heaps_equal

    x = synthetic.mixture(n, dim, 32, 0.3, 500 + dim, 501 + n)
    x[50:90] = x[49]
    q = synthetic.mixture(7, dim, 32, 0.3, 500 + dim, 502)
    packed, alpha = oracle.encode(x, bits)
    half = n // 2 + 17
    hg, ho = empty_heaps(7, k, {6: 0.1}), empty_heaps(7, k, {6: 0.1})
    for ci, (a, b) in enumerate(((0, half), (half, n))):
        perm0 = oracle.repack_perm0(packed[a:b], bits, dim)
        qidx = np.array([6, 2, 0, 5] if ci == 0 else [1, 2, 3, 4, 5, 6, 0], np.uint64)
        off = np.linspace(-1, 1, 7).astype(np.float32) if ci else None
        idm = (np.arange(b - a, dtype=np.uint64) * 3 + (7 if ci else 100000))
        ho = oracle.score_into(q, k, bits, perm0, alpha[a:b], qidx, ho, id_map=idm, offsets=off)
        vc.score_perm0_into(q, k, bits, perm0, alpha[a:b], qidx, hg, id_map=idm, score_offsets=off)
        assert heaps_equal(hg, ho), (bits, dim, k, ci)


# ------------------------------------------------------ prepared queries
def test_prepared_queries_match_golden(vc, golden_cases):
    """prepare_queries kept on the device: its state equals the reference's
    (LUT bytes, scales, biases, bias_corr); prepared_query_at(i) holds row i;
    search through the handle (score_prepared over the index's PERM0 export)
    equals the reference's top-k; repacking the golden bit-planes equals the
    golden flat bytes, PERM0 and sequential blocks."""
    for c, g, x, q, sample in golden_cases:
        sh, sc = (g["shift"], g["scale"]) if c["tq"] else (None, None)
        prep = vc.prepare_queries(q, c["bits"], sh, sc)
        st = prep.export()
        assert np.array_equal(st["luts"], g["prep_luts"]), c
        assert bits_equal(st["lut_scale"], g["prep_lut_scale"]), c
        assert bits_equal(st["lut_bias"], g["prep_lut_bias"]), c
        assert bits_equal(st["bias_corr"], g["prep_bias_corr"]), c
        one = prep.query_at(3).export()
        assert np.array_equal(one["luts"][0], g["prep_luts"][3]) and bits_equal(one["bias_corr"],
                                                                                g["prep_bias_corr"][3:4]), c
        n, dim, bits = c["n"], c["dim"], c["bits"]
        assert np.array_equal(vc.codes_flat_from_packed(g["packed"], n, dim, bits), g["flat"]), c
        perm0 = vc.repack_perm0(g["packed"], n, dim, bits)
        assert sha(perm0) == bytes(g["perm0_sha"]).hex(), c
        assert sha(vc.repack_seq(g["packed"], n, dim, bits)) == bytes(g["seq_sha"]).hex(), c
        for k in (1, 10, 64):
            r = vc.score_prepared(prep, k, perm0, g["alpha"])
            assert np.array_equal(r.ids, g[f"ids_k{k}"]), (c, k)
            assert scores_equal(r.scores, g[f"scores_k{k}"]), (c, k)
        prep.close()


def test_prepared_score_arguments_match_golden(vc, oracle):
    """score_prepared through one prepared handle per case, every argument
    (id_map, query subsets with a repeat, residual correction), against the
    reference build's golden outputs."""
    from cases import SCORE_PREPARED_CASES, SCORE_PREPARED_VARIANTS, score_prepared_inputs
    from conftest import load_golden

    g = load_golden("score_prepared.npz")
    for ci, (bits, dim, n, nq, k) in enumerate(SCORE_PREPARED_CASES):
        x, q, *extra = score_prepared_inputs(bits, dim, n, nq)
        packed, alpha = oracle.encode(x, bits)
        perm0 = oracle.repack_perm0(packed, bits, dim)
        prep = vc.prepare_queries(q, bits)
        for tag, kw in SCORE_PREPARED_VARIANTS(*extra):
            r = vc.score_prepared(prep, k, perm0, alpha, **kw)
            assert np.array_equal(r.ids, g[f"c{ci}_{tag}_ids"]), (ci, tag)
            assert bits_equal(r.scores, g[f"c{ci}_{tag}_scores"]), (ci, tag)
        with pytest.raises(IndexError):
            vc.score_prepared(prep, k, perm0, alpha, query_indices=np.array([nq], np.uint64))
        with pytest.raises(IndexError):
            prep.query_at(nq)
        prep.close()


def test_prepared_into_chain_matches_golden(vc):
    """The score_prepared_into golden chains with the queries prepared once
    and the handle reused by every call: every heap array after every call
    equals the reference's."""
    from conftest import load_golden
    from util import empty_heaps, heaps_equal, score_into_scenarios

    g = load_golden("score_into.npz")
    checked = 0
    for name, bits, dim, k, q, shift, scale, calls, want in score_into_scenarios(g):
        prep = vc.prepare_queries(q, bits, shift, scale)
        heaps = empty_heaps(q.shape[0], k, {4: 0.05})
        for ci, (perm0, alpha, id_map, qidx, off) in enumerate(calls):
            vc.score_prepared_into(prep, k, perm0, alpha, qidx, heaps, id_map=id_map, score_offsets=off)
            assert heaps_equal(heaps, want[ci]), (name, k, ci)
            checked += 1
        prep.close()
    assert checked >= 10


def test_repack_matches_oracle(vc, oracle):
    """The repack free functions on the GPU equal the oracle's on ragged sizes
    (a partial last block of 192) for every bit width."""
    from vectorcache_amd import synthetic

    for bits, dim, n in ((4, 128, 1000), (3, 200, 385), (2, 768, 577), (4, 1536, 193)):
        x = synthetic.mixture(n, dim, 16, 0.3, 900 + dim, 901 + n)
        packed, _ = oracle.encode(x, bits)
        assert np.array_equal(vc.codes_flat_from_packed(packed, n, dim, bits),
                              oracle.codes_flat(packed, bits, dim).reshape(n, -1)), (bits, dim)
        assert np.array_equal(vc.repack_perm0(packed, n, dim, bits), oracle.repack_perm0(packed, bits, dim).ravel()), \
            (bits, dim)
        assert np.array_equal(vc.repack_seq(packed, n, dim, bits), oracle.repack_seq(packed, bits, dim).ravel()), \
            (bits, dim)


# ------------------------------------------------------------ float re-rank
def test_float_rescore_matches_golden(vc):
    """float_rescore_topk on the GPU, host buffers: every dim's rows equal
    the reference's bit for bit (all entries, and with per-row counts)."""
    from cases import RESCORE_DIMS, rescore_inputs
    from conftest import load_golden

    g = load_golden("float_rescore.npz")
    for dim in RESCORE_DIMS:
        corpus, q, scores, ids, counts = rescore_inputs(dim)
        for tag, cnt in (("all", None), ("counts", counts)):
            r = vc.float_rescore_topk(vc.SearchResults(scores, ids), q, corpus, cnt)
            assert bits_equal(r.scores, g[f"d{dim}_{tag}_scores"]) and np.array_equal(r.ids, g[f"d{dim}_{tag}_ids"]), \
                (dim, tag)


def test_float_rescore_device_after_search(vc, oracle):
    """Search, then re-rank the device-resident top-40 against the device
    corpus (torch tensors, in place): equals the oracle's re-rank of the same
    rows, and the re-ranked top-10 agrees with the FP64 exact top-10 at least
    as often as the LUT top-10 does."""
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.mixture(20000, 768, 64, 0.3, 61, 62)
    q = synthetic.mixture(32, 768, 64, 0.3, 61, 63)
    ix = vc.TurboQuantIndex(768, 4)
    ix.add(x)
    ix.prepare()
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x).to(dev)
    qd = torch.from_numpy(q).to(dev)
    s, i = ix.search_device(qd, 40)
    torch.cuda.synchronize()
    s0, i0 = s.cpu().numpy().copy(), i.cpu().numpy().astype(np.uint64)
    r = vc.float_rescore_topk(vc.SearchResults(s, i), qd, xd)
    torch.cuda.synchronize()
    so, io = oracle.float_rescore(s0, i0, q, x)
    assert bits_equal(r.scores.cpu().numpy(), so) and np.array_equal(r.ids.cpu().numpy().astype(np.uint64), io)
    exact = np.argsort(-(q.astype(np.float64) @ x.astype(np.float64).T), axis=1, kind="stable")[:, :10]
    rec_lut = np.mean([len(set(i0[j, :10]) & set(exact[j])) for j in range(32)])
    rec_rr = np.mean([len(set(io[j, :10]) & set(exact[j])) for j in range(32)])
    assert rec_rr >= rec_lut


def test_search_device_orders_on_current_stream(vc):
    """search_device defaults to torch's current stream: fp16, non-contiguous
    queries made on that stream (no synchronize before the call) give the
    results of a host search of the same rounded queries; CPU tensors, rows
    that are not whole and short or mistyped output buffers are rejected."""
    import torch

    from vectorcache_amd import synthetic

    x = synthetic.mixture(6000, 128, 32, 0.3, 71, 72)
    q = synthetic.mixture(24, 128, 32, 0.3, 71, 73)
    q16 = q.astype(np.float16).astype(np.float32)
    ix = vc.TurboQuantIndex(128, 4)
    ix.add(x)
    ix.prepare()
    want = ix.search(q16, 10)
    dev = torch.device("cuda:0")
    torch.cuda.synchronize()
    # a large producer kernel right before the call: the search must wait for it
    big = torch.randn((4096, 4096), device=dev)
    qt = (torch.from_numpy(np.ascontiguousarray(q.T)).to(dev).half() + big[:128, :24].half() * 0)
    s, i = ix.search_device(qt.t(), 10)
    got_s, got_i = s.cpu().numpy(), i.cpu().numpy().astype(np.uint64)
    assert np.array_equal(got_i, want.ids) and scores_equal(got_s, want.scores)
    with pytest.raises(ValueError):
        ix.search_device(torch.from_numpy(q), 10)
    with pytest.raises(ValueError):
        ix.search_device(torch.from_numpy(q).to(dev)[:, :100], 10)
    with pytest.raises(ValueError):
        ix.search_device(torch.from_numpy(q).to(dev), 10, scores=torch.empty((24, 5), device=dev))
    with pytest.raises(ValueError):
        ix.search_device(torch.from_numpy(q).to(dev), 10, ids=torch.empty((24, 10), dtype=torch.int32, device=dev))
