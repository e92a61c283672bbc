"""Test infrastructure only: ctypes bindings for the two CPU checkers.

* ``Oracle("port")``      -> oracle/liboracle.so, this repository's C restatement
  of the reference's plain flat TurboQuant path (oracle/oracle.c);
* ``Oracle("reference")`` -> oracle/_ref/libvcref.so, the reference's own
  sources compiled in place (oracle/Makefile, oracle/ref_harness.cpp).

Both expose the same numpy-level functions. Only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg import this module, as the checker; the product
path (vectorcache_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_P = ctypes.c_void_p
_S = ctypes.c_size_t


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


class Oracle:
    def __init__(self, kind: str = "port"):
        self.kind = kind
        if kind == "port":
            path, self._pre = HERE / "liboracle.so", "oc_"
        elif kind == "reference":
            path, self._pre = HERE / "_ref" / "libvcref.so", "vcref_"
        else:
            raise ValueError(kind)
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (make -C oracle{' ref' if kind == 'reference' else ''})")
        if kind == "reference":
            os.environ["TURBOVEC_NO_VNNI"] = "1"
        self.lib = ctypes.CDLL(str(path))
        # attribute access caches the function object, so restype sticks
        self._f("row_norm").restype = ctypes.c_float
        self._f("beta_cdf_aa").restype = ctypes.c_double

    def _f(self, name):
        return getattr(self.lib, self._pre + name)

    def _check(self, rc, name):
        if rc != 0:
            msg = ""
            if self.kind == "reference":
                f = self.lib.vcref_last_error
                f.restype = ctypes.c_char_p
                msg = f().decode()
            raise RuntimeError(f"{self.kind}:{name} failed ({rc}) {msg}")

    def set_threads(self, n: int):
        self._f("set_threads")(ctypes.c_int(int(n)))

    # ---- host precompute -------------------------------------------------
    def rotation_tables(self, dim):
        perm = np.zeros(dim, np.uint32)
        signs = np.zeros(dim, np.float32)
        isb, blk = ctypes.c_float(), ctypes.c_size_t()
        self._check(self._f("rotation_tables")(_S(dim), _ptr(perm), _ptr(signs), ctypes.byref(isb),
                                               ctypes.byref(blk)), "rotation_tables")
        return perm, signs, np.float32(isb.value), int(blk.value)

    def codebook(self, bits, dim):
        b = np.zeros(15, np.float32)
        c = np.zeros(16, np.float32)
        self._check(self._f("codebook")(_S(bits), _S(dim), _ptr(b), _ptr(c)), "codebook")
        L = 1 << bits
        return b[: L - 1].copy(), c[:L].copy()

    def beta_cdf_aa(self, a, t):
        return float(self._f("beta_cdf_aa")(ctypes.c_double(a), ctypes.c_double(t)))

    def row_norm(self, row):
        row = np.ascontiguousarray(row, np.float32)
        return np.float32(self._f("row_norm")(_ptr(row), _S(row.size)))

    def rotate_scaled(self, src, inv):
        src = np.ascontiguousarray(src, np.float32)
        dst = np.zeros_like(src)
        self._check(self._f("rotate_scaled")(_S(src.size), _ptr(src), ctypes.c_float(inv), _ptr(dst)),
                    "rotate_scaled")
        return dst

    # ---- encode -----------------------------------------------------------
    def encode(self, x, bits, shift=None, scale=None, want_rotated=False):
        x = np.ascontiguousarray(x, np.float32)
        n, dim = x.shape
        packed = np.zeros((n, bits * dim // 8), np.uint8)
        alpha = np.zeros(n, np.float32)
        rot = np.zeros((n, dim), np.float32) if want_rotated else None
        sh = None if shift is None else np.ascontiguousarray(shift, np.float32)
        sc = None if scale is None else np.ascontiguousarray(scale, np.float32)
        self._check(self._f("encode")(_ptr(x), _S(n), _S(dim), _S(bits), _ptr(sh), _ptr(sc),
                                      _ptr(packed), _ptr(alpha), _ptr(rot)), "encode")
        return (packed, alpha, rot) if want_rotated else (packed, alpha)

    def fit_calibration(self, x, bits):
        x = np.ascontiguousarray(x, np.float32)
        n, dim = x.shape
        shift = np.zeros(dim, np.float32)
        scale = np.zeros(dim, np.float32)
        self._check(self._f("fit_calibration")(_ptr(x), _S(n), _S(dim), _S(bits), _ptr(shift),
                                               _ptr(scale)), "fit_calibration")
        return shift, scale

    # ---- pack -------------------------------------------------------------
    def codes_flat(self, packed, bits, dim):
        packed = np.ascontiguousarray(packed, np.uint8)
        n = packed.shape[0]
        G = dim // (8 // bits)
        out = np.zeros((n, G), np.uint8)
        self._check(self._f("codes_flat")(_ptr(packed), _S(n), _S(bits), _S(dim), _ptr(out)),
                    "codes_flat")
        return out

    def repack_perm0(self, packed, bits, dim):
        """bit-planes -> PERM0 192x1 blocked bytes."""
        packed = np.ascontiguousarray(packed, np.uint8)
        n = packed.shape[0]
        G = dim // (8 // bits)
        nb = (n + 191) // 192
        out = np.zeros(nb * G * 192, np.uint8)
        if self.kind == "reference":
            nbo = ctypes.c_size_t()
            self._check(self.lib.vcref_repack_perm0(_ptr(packed), _S(n), _S(bits), _S(dim), _ptr(out),
                                                    ctypes.byref(nbo)), "repack_perm0")
        else:
            flat = self.codes_flat(packed, bits, dim)
            self._check(self.lib.oc_pack_perm0(_ptr(flat), _S(n), _S(G), _ptr(out)), "pack_perm0")
        return out

    def flat_to_gpu32(self, flat):
        """Row-major byte groups (n x G) -> the reference CUDA index's u32
        words [ceil(n/32)][ceil(G/4)][32]."""
        flat = np.ascontiguousarray(flat, np.uint8)
        n, G = flat.shape
        out = np.zeros(((n + 31) // 32) * ((G + 3) // 4) * 32, np.uint32)
        self._check(self.lib.oc_flat_to_gpu32(_ptr(flat), _S(n), _S(G), _ptr(out)), "flat_to_gpu32")
        return out

    def gpu32_to_flat(self, words, n, G):
        words = np.ascontiguousarray(words, np.uint32)
        out = np.zeros((n, G), np.uint8)
        self._check(self.lib.oc_gpu32_to_flat(_ptr(words), _S(n), _S(G), _ptr(out)), "gpu32_to_flat")
        return out

    def repack_seq(self, packed, bits, dim):
        packed = np.ascontiguousarray(packed, np.uint8)
        n = packed.shape[0]
        G = dim // (8 // bits)
        nb = (n + 191) // 192
        out = np.zeros(nb * G * 192, np.uint8)
        if self.kind == "reference":
            self._check(self.lib.vcref_repack_seq(_ptr(packed), _S(n), _S(bits), _S(dim), _ptr(out)),
                        "repack_seq")
        else:
            flat = self.codes_flat(packed, bits, dim)
            self._check(self.lib.oc_pack_seq(_ptr(flat), _S(n), _S(G), _ptr(out)), "pack_seq")
        return out

    def unpack_perm0(self, blocked, n, G):
        """PERM0 blocked -> row-major flat byte groups (port only)."""
        blocked = np.ascontiguousarray(blocked, np.uint8)
        out = np.zeros((n, G), np.uint8)
        self._check(self.lib.oc_unpack_perm0(_ptr(blocked), _S(n), _S(G), _ptr(out)), "unpack_perm0")
        return out

    # ---- search -------------------------------------------------------------
    def prepare_queries(self, q, bits, shift=None, scale=None):
        q = np.ascontiguousarray(q, np.float32)
        nq, dim = q.shape
        G = dim // (8 // bits)
        q_rot = np.zeros((nq, dim), np.float32)
        bc = np.zeros(nq, np.float32)
        luts = np.zeros((nq, G * 32), np.uint8)
        ls = np.zeros(nq, np.float32)
        lb = np.zeros(nq, np.float32)
        sh = None if shift is None else np.ascontiguousarray(shift, np.float32)
        sc = None if scale is None else np.ascontiguousarray(scale, np.float32)
        self._check(self._f("prepare_queries")(_ptr(q), _S(nq), _S(dim), _S(bits), _ptr(sh), _ptr(sc),
                                               _ptr(q_rot), _ptr(bc), _ptr(luts), _ptr(ls), _ptr(lb)),
                    "prepare_queries")
        return dict(q_rot=q_rot, bias_corr=bc, luts=luts, lut_scale=ls, lut_bias=lb)

    def search_flat(self, q, k, bits, blocked_perm0, alphas, shift=None, scale=None):
        q = np.ascontiguousarray(q, np.float32)
        nq, dim = q.shape
        n = alphas.size
        alphas = np.ascontiguousarray(alphas, np.float32)
        blocked_perm0 = np.ascontiguousarray(blocked_perm0, np.uint8)
        kk = max(min(k, n), 0)
        scores = np.zeros(max(nq * kk, 1), np.float32)
        ids = np.zeros(max(nq * kk, 1), np.uint64)
        kout = ctypes.c_size_t()
        sh = None if shift is None else np.ascontiguousarray(shift, np.float32)
        sc = None if scale is None else np.ascontiguousarray(scale, np.float32)
        if self.kind == "reference":
            nb = (n + 191) // 192
            rc = self.lib.vcref_search_flat(_ptr(q), _S(nq), _S(dim), _S(k), _S(bits), _ptr(blocked_perm0),
                                            _S(nb), _ptr(alphas), _S(n), _ptr(sh), _ptr(sc), _ptr(scores),
                                            _ptr(ids), ctypes.byref(kout))
        else:
            rc = self.lib.oc_search_flat(_ptr(q), _S(nq), _S(dim), _S(k), _S(bits), _ptr(blocked_perm0),
                                         _ptr(alphas), _S(n), _ptr(sh), _ptr(sc), _ptr(scores), _ptr(ids),
                                         ctypes.byref(kout))
        self._check(rc, "search_flat")
        ko = int(kout.value)
        return scores[: nq * ko].reshape(nq, ko), ids[: nq * ko].reshape(nq, ko)

    def score_prepared(self, q, k, bits, blocked_perm0, alphas, shift=None, scale=None, id_map=None,
                       query_indices=None, residual_norms=None, residual_scales=None):
        """score_prepared with id_map, query_indices and the residual
        post-correction (AVX2 path). Returns (scores, ids) nq_out x k_eff."""
        q = np.ascontiguousarray(q, np.float32)
        nq, dim = q.shape
        alphas = np.ascontiguousarray(alphas, np.float32)
        n = alphas.size
        blocked_perm0 = np.ascontiguousarray(blocked_perm0, np.uint8)
        qi = None if query_indices is None else np.ascontiguousarray(query_indices, np.uint64)
        nq_out = nq if qi is None else qi.size
        kk = max(min(k, n), 0)
        scores = np.zeros(max(nq_out * kk, 1), np.float32)
        ids = np.zeros(max(nq_out * kk, 1), np.uint64)
        kout = ctypes.c_size_t()
        cv = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
        sh, sc, im = cv(shift, np.float32), cv(scale, np.float32), cv(id_map, np.uint64)
        rn, rs = cv(residual_norms, np.float32), cv(residual_scales, np.float32)
        args = (_ptr(im), _ptr(qi), _S(0 if qi is None else qi.size), _ptr(rn), _ptr(rs), _ptr(scores), _ptr(ids),
                ctypes.byref(kout))
        if self.kind == "reference":
            nb = (n + 191) // 192
            rc = self.lib.vcref_score_prepared(_ptr(q), _S(nq), _S(dim), _S(k), _S(bits), _ptr(blocked_perm0),
                                               _S(nb), _ptr(alphas), _S(n), _ptr(sh), _ptr(sc), *args)
        else:
            rc = self.lib.oc_score_prepared(_ptr(q), _S(nq), _S(dim), _S(k), _S(bits), _ptr(blocked_perm0),
                                            _ptr(alphas), _S(n), _ptr(sh), _ptr(sc), *args)
        self._check(rc, "score_prepared")
        ko = int(kout.value)
        return scores[: nq_out * ko].reshape(nq_out, ko), ids[: nq_out * ko].reshape(nq_out, ko)

    def scan_topk(self, luts, lut_scale, lut_bias_total, flat, alphas, k, id_map=None):
        """Scan + top-k from prepared LUTs and row-major flat codes (port only)."""
        luts = np.ascontiguousarray(luts, np.uint8)
        flat = np.ascontiguousarray(flat, np.uint8)
        nq = luts.shape[0]
        n, G = flat.shape
        kk = max(min(k, n), 0)
        scores = np.zeros(max(nq * kk, 1), np.float32)
        ids = np.zeros(max(nq * kk, 1), np.uint64)
        kout = ctypes.c_size_t()
        im = None if id_map is None else np.ascontiguousarray(id_map, np.uint64)
        self._check(self.lib.oc_scan_topk(_ptr(luts), _ptr(np.ascontiguousarray(lut_scale, np.float32)),
                                          _ptr(np.ascontiguousarray(lut_bias_total, np.float32)), _S(nq),
                                          _ptr(flat), _S(n), _S(G), _ptr(np.ascontiguousarray(alphas, np.float32)),
                                          _ptr(im), _S(k), _ptr(scores), _ptr(ids), ctypes.byref(kout)),
                    "scan_topk")
        ko = int(kout.value)
        return scores[: nq * ko].reshape(nq, ko), ids[: nq * ko].reshape(nq, ko)

    def window_sums(self, lut, flat):
        lut = np.ascontiguousarray(lut, np.uint8)
        flat = np.ascontiguousarray(flat, np.uint8)
        n, G = flat.shape
        W = (G + 255) // 256
        out = np.zeros((n, W), np.uint32)
        self._check(self.lib.oc_window_sums(_ptr(lut), _ptr(flat), _S(n), _S(G), _ptr(out)), "window_sums")
        return out

    def merge_results(self, dst_scores, dst_ids, src_scores, src_ids):
        ds = np.ascontiguousarray(dst_scores, np.float32).copy()
        di = np.ascontiguousarray(dst_ids, np.uint64).copy()
        ss = np.ascontiguousarray(src_scores, np.float32)
        si = np.ascontiguousarray(src_ids, np.uint64)
        nq, kd = ds.shape
        ks = ss.shape[1]
        self._check(self._f("merge_results")(_S(nq), _S(kd), _ptr(ds), _ptr(di), _S(ks), _ptr(ss), _ptr(si)),
                    "merge_results")
        return ds, di

    def score_into(self, q, k, bits, blocked_perm0, alphas, qidx, heaps, id_map=None, shift=None, scale=None,
                   offsets=None):
        """score_prepared_into over one bucket (PERM0 blocks + alphas): heaps is
        (heap_s [nq, k] f32, heap_i [nq, k] u64, heap_sz [nq] u64, heap_min [nq]
        f32, heap_mi [nq] u64); returns the updated copies."""
        q = np.ascontiguousarray(q, np.float32)
        nq, dim = q.shape
        alphas = np.ascontiguousarray(alphas, np.float32)
        n = alphas.size
        blocked_perm0 = np.ascontiguousarray(blocked_perm0, np.uint8)
        qidx = np.ascontiguousarray(qidx, np.uint64)
        hs, hi, hsz, hmin, hmi = (np.ascontiguousarray(heaps[0], np.float32).copy(),
                                  np.ascontiguousarray(heaps[1], np.uint64).copy(),
                                  np.ascontiguousarray(heaps[2], np.uint64).copy(),
                                  np.ascontiguousarray(heaps[3], np.float32).copy(),
                                  np.ascontiguousarray(heaps[4], np.uint64).copy())
        im = None if id_map is None else np.ascontiguousarray(id_map, np.uint64)
        sh = None if shift is None else np.ascontiguousarray(shift, np.float32)
        sc = None if scale is None else np.ascontiguousarray(scale, np.float32)
        off = None if offsets is None else np.ascontiguousarray(offsets, np.float32)
        if self.kind == "reference":
            nb = (n + 191) // 192
            rc = self.lib.vcref_score_into(_ptr(q), _S(nq), _S(dim), _S(k), _S(bits), _ptr(blocked_perm0), _S(nb),
                                           _ptr(alphas), _S(n), _ptr(im), _ptr(sh), _ptr(sc), _ptr(qidx),
                                           _S(qidx.size), _ptr(hs), _ptr(hi), _ptr(hsz), _ptr(hmin), _ptr(hmi),
                                           _ptr(off))
        else:
            rc = self.lib.oc_score_into(_ptr(q), _S(nq), _S(dim), _S(k), _S(bits), _ptr(blocked_perm0),
                                        _ptr(alphas), _S(n), _ptr(im), _ptr(sh), _ptr(sc), _ptr(qidx),
                                        _S(qidx.size), _ptr(hs), _ptr(hi), _ptr(hsz), _ptr(hmin), _ptr(hmi),
                                        _ptr(off))
        self._check(rc, "score_into")
        return hs, hi, hsz, hmin, hmi

    # ---- bucketed index (its scores; bucket choice is not part of the oracle) ----
    def bucket_residuals(self, x, pos_rows, pos_bucket, cent):
        """Rows x_hat - c of the given positions (cent None: x_hat)."""
        x = np.ascontiguousarray(x, np.float32)
        dim = x.shape[1]
        pr = np.ascontiguousarray(pos_rows, np.uint64)
        pbk = np.ascontiguousarray(pos_bucket, np.uint32)
        c = None if cent is None else np.ascontiguousarray(cent, np.float32)
        out = np.zeros((pr.size, dim), np.float32)
        self._check(self._port("bucket_residuals")(_ptr(x), _S(dim), _ptr(pr), _ptr(pbk), _S(pr.size), _ptr(c), _ptr(out)),
                    "bucket_residuals")
        return out

    def probe_offsets(self, q, cent, probe_b, probe_n):
        q = np.ascontiguousarray(q, np.float32)
        nq, dim = q.shape
        pb = np.ascontiguousarray(probe_b, np.int32)
        pn = np.ascontiguousarray(probe_n, np.int32)
        c = np.ascontiguousarray(cent, np.float32)
        off = np.zeros(pb.shape, np.float32)
        self._check(self._port("probe_offsets")(_ptr(q), _S(nq), _S(dim), _ptr(c), _ptr(pb), _S(pb.shape[1]), _ptr(pn),
                                             _ptr(off)), "probe_offsets")
        return off

    def _port(self, name):
        """A function of this repository's C restatement only (the reference
        has no code for the bucketed index)."""
        if self.kind != "port":
            raise ValueError(f"{name} exists in the port oracle only")
        return getattr(self.lib, "oc_" + name)

    def bucketed_search(self, q, k, bits, bucket_sizes, flat, alpha, pos_ids, cent, probe_b, probe_n, shift=None,
                        scale=None, residual=True):
        """The bucketed index's search as a chain of score_prepared_into calls
        (one per bucket, every query that probes it, offsets <q_hat, c>), then
        heaps_to_search_results. probe_b / probe_n: each query's probed
        buckets (the routing is not part of the oracle)."""
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        sizes = np.asarray(bucket_sizes, np.int64)
        boff = np.concatenate([[0], np.cumsum(sizes)])
        G = flat.shape[1]
        pb = np.asarray(probe_b, np.int32)
        pn = np.asarray(probe_n, np.int32)
        off = self.probe_offsets(q, cent, pb, pn) if residual else np.zeros(pb.shape, np.float32)
        heaps = (np.zeros((nq, k), np.float32), np.zeros((nq, k), np.uint64), np.zeros(nq, np.uint64),
                 np.full(nq, -np.inf, np.float32), np.zeros(nq, np.uint64))
        users = {}
        for i in range(nq):
            for j in range(int(pn[i])):
                users.setdefault(int(pb[i, j]), []).append((i, j))
        for b in sorted(users):
            n = int(sizes[b])
            if n == 0:
                continue
            rows = slice(int(boff[b]), int(boff[b]) + n)
            fl = np.ascontiguousarray(flat[rows])
            perm0 = np.zeros(((n + 191) // 192) * G * 192, np.uint8)
            self._check(self._port("pack_perm0")(_ptr(fl), _S(n), _S(G), _ptr(perm0)), "pack_perm0")
            qi = np.array([i for i, _ in users[b]], np.uint64)
            offs = np.zeros(nq, np.float32)
            for i, j in users[b]:
                offs[i] = off[i, j]
            heaps = self.score_into(q, k, bits, perm0, alpha[rows], qi, heaps, id_map=pos_ids[rows], shift=shift,
                                    scale=scale, offsets=offs)
        return self.heaps_to_results(heaps[0], heaps[1], heaps[2])

    def heaps_to_results(self, heap_s, heap_i, heap_sz):
        hs = np.ascontiguousarray(heap_s, np.float32)
        hi = np.ascontiguousarray(heap_i, np.uint64)
        hsz = np.ascontiguousarray(heap_sz, np.uint64)
        nq, k = hs.shape
        scores = np.zeros((nq, k), np.float32)
        ids = np.zeros((nq, k), np.uint64)
        self._check(self._f("heaps_to_results")(_S(nq), _S(k), _ptr(hs), _ptr(hi), _ptr(hsz), _ptr(scores),
                                                _ptr(ids)), "heaps_to_results")
        return scores, ids

    def float_rescore(self, scores, ids, queries, corpus, counts=None):
        """float_rescore_topk: returns re-scored, re-ordered copies of the rows."""
        sc = np.ascontiguousarray(scores, np.float32).copy()
        ii = np.ascontiguousarray(ids, np.uint64).copy()
        nq, k = sc.shape
        q = np.ascontiguousarray(queries, np.float32)
        cor = np.ascontiguousarray(corpus, np.float32)
        dim = q.shape[1]
        cnt = None if counts is None else np.ascontiguousarray(counts, np.uint64)
        self._check(self._f("float_rescore")(_S(nq), _S(k), _ptr(sc), _ptr(ii), _ptr(cnt), _ptr(q), _ptr(cor),
                                             _S(cor.shape[0]), _S(dim)), "float_rescore")
        return sc, ii

    # ---- landmark index ----------------------------------------------------
    def ortho_landmarks(self, dim, m, seed):
        L = np.zeros(dim * m, np.float32)
        self._check(self._f("ortho_landmarks")(_S(dim), _S(m), ctypes.c_uint64(seed), _ptr(L)), "ortho_landmarks")
        return L.reshape(m, dim)

    def quantize_landmarks(self, L):
        L = np.ascontiguousarray(L, np.float32)
        m, dim = L.shape
        codes = np.zeros((m, ((dim + 15) // 16 * 16) // 2), np.uint8)
        scales = np.zeros(m, np.float32)
        self._check(self._f("quantize_landmarks")(_ptr(L), _S(dim), _S(m), _ptr(codes), _ptr(scales)),
                    "quantize_landmarks")
        return codes, scales

    def project_rows(self, Y, L, stride):
        Y = np.ascontiguousarray(Y, np.float32)
        L = np.ascontiguousarray(L, np.float32)
        out = np.zeros((Y.shape[0], stride), np.float32)
        self._check(self._f("project_rows")(_ptr(Y), _S(Y.shape[0]), _S(Y.shape[1]), _ptr(L), _S(L.shape[0]),
                                            _ptr(out), _S(stride)), "project_rows")
        return out

    def project_q4(self, q_rot, L, stride):
        q = np.ascontiguousarray(q_rot, np.float32)
        L = np.ascontiguousarray(L, np.float32)
        out = np.zeros((q.shape[0], stride), np.float32)
        self._check(self._f("project_q4")(_ptr(q), _S(q.shape[0]), _S(q.shape[1]), _ptr(L), _S(L.shape[0]),
                                          _ptr(out), _S(stride)), "project_q4")
        return out

    def lm_search(self, x, q, k, scan_bits=4, n_landmarks=1024, seed=0, rerank_bits=0, batches=None,
                  prepare_between=False, float_corpus=None, planes=False):
        """LandmarkIndex(dim, scan_bits, n_landmarks, seed, rerank_bits): add the
        rows of x (in `batches` row counts), prepare, search. Returns scores,
        ids, info {scan_dim, codes_bytes, index_bytes, search_s (reference: wall time
        of search())}. planes=True: as on an AVX-512 VNNI host (the staged
        sign-plane branch; the reference build runs with its VNNI kernels)."""
        x = np.ascontiguousarray(x, np.float32)
        q = np.ascontiguousarray(q, np.float32)
        n, dim = x.shape
        nq = q.shape[0]
        br = np.asarray(batches if batches is not None else [n], np.uint64)
        scores = np.zeros(max(1, nq * k), np.float32)
        ids = np.zeros(max(1, nq * k), np.uint64)
        ko = ctypes.c_size_t()
        info = np.zeros(3, np.uint64)
        secs = ctypes.c_double()
        fc = None if float_corpus is None else np.ascontiguousarray(float_corpus, np.float32)
        fn = "lm_search_planes" if planes else "lm_search"
        self._check(self._f(fn)(_ptr(x), _ptr(br), _S(len(br)), ctypes.c_int(int(prepare_between)),
                                         _ptr(q), _S(nq), _S(dim), _S(scan_bits), _S(n_landmarks),
                                         ctypes.c_uint64(seed), _S(rerank_bits), _S(k), _ptr(fc), _ptr(scores),
                                         _ptr(ids), ctypes.byref(ko), _ptr(info), ctypes.byref(secs)), fn)
        ke = ko.value
        return (scores[: nq * ke].reshape(nq, ke), ids[: nq * ke].reshape(nq, ke),
                {"scan_dim": int(info[0]), "codes_bytes": int(info[1]), "index_bytes": int(info[2]),
                 "search_s": secs.value})

    def sign_topk(self, rows, packed, k):
        """The sign-plane scan alone: rows (nq x S), bit-plane codes (n x 4 S/8)
        -> top k (scores, ids) by (score desc, id asc)."""
        rows = np.ascontiguousarray(rows, np.float32)
        packed = np.ascontiguousarray(packed, np.uint8)
        nq, S = rows.shape
        n = packed.shape[0]
        scores = np.zeros(max(1, nq * k), np.float32)
        ids = np.zeros(max(1, nq * k), np.uint64)
        ko = ctypes.c_size_t()
        self._check(self._f("sign_topk")(_ptr(rows), _S(nq), _S(S), _ptr(packed), _S(n), _S(k), _ptr(scores),
                                         _ptr(ids), ctypes.byref(ko)), "sign_topk")
        ke = ko.value
        return scores[: nq * ke].reshape(nq, ke), ids[: nq * ke].reshape(nq, ke)

    def set_staged_variant(self, v: int):
        """The oracle's fusion choices in the staged branch (test hook: 0 is the
        reference's)."""
        self.lib.oc_set_staged_variant(ctypes.c_int(v))

    def planes_exact_len(self, k: int) -> int:
        f = self.lib.oc_planes_exact_len
        f.restype = ctypes.c_size_t
        return int(f(_S(k)))


def available(kind: str) -> bool:
    if kind == "port":
        return (HERE / "liboracle.so").exists()
    return (HERE / "_ref" / "libvcref.so").exists()
