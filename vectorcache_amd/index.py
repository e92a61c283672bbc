"""TurboQuantIndex: the plain flat TurboQuant index on the GPU.

Mirrors the reference's index class shape (include/vectorcache/index.hpp:18-78)
and the README's calibrate()/add()/prepare()/search() flow for the plain flat
path; every call goes through the C ABI of libvectorcache_hip.so. Errors map
onto the reference's exception types: std::invalid_argument -> ValueError,
std::runtime_error (call order) -> IndexStateError, std::out_of_range ->
IndexError.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import P, S, check, lib


def _f32(a, cols=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if cols is not None:
        if a.size % cols != 0:
            raise ValueError(f"input size {a.size} is not a multiple of dim {cols}")
        a = a.reshape(-1, cols)
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def _cuda_rows(t, dim: int, name: str):
    """A CUDA tensor of whole rows of `dim` as contiguous float32 (a copy made
    on torch's current stream when it is not one already)."""
    if not (_is_torch(t) and t.is_cuda):
        raise ValueError(f"{name} must be a CUDA tensor")
    t = t.contiguous().float()
    if t.numel() % dim != 0:
        raise ValueError(f"{name} size {t.numel()} is not a multiple of dim {dim}")
    return t


def _cuda_out(buf, rows: int, cols: int, dtype, name: str, device):
    """A caller output buffer checked for device, dtype, layout and room for
    rows x cols, or a new one."""
    import torch

    if buf is None:
        return torch.empty((rows, cols), dtype=dtype, device=device)
    if not (_is_torch(buf) and buf.is_cuda):
        raise ValueError(f"{name} must be a CUDA tensor")
    if buf.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {buf.dtype}")
    if not buf.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if buf.numel() < rows * cols:
        raise ValueError(f"{name} holds {buf.numel()} elements, {rows} x {cols} needed")
    return buf


def _launch_stream(stream, *tensors):
    """The HIP stream a device call is ordered on: torch's current stream by
    default, so the kernels run after the work that produced the inputs and
    before whatever torch enqueues next. A caller-given stream also keeps the
    temporaries alive for it."""
    import torch

    if stream is None:
        return P(torch.cuda.current_stream().cuda_stream)
    cur = torch.cuda.current_stream().cuda_stream
    if int(stream) != int(cur):
        ext = torch.cuda.ExternalStream(int(stream))
        for t in tensors:
            if t is not None:
                t.record_stream(ext)
    return P(int(stream))


@dataclass
class SearchResults:
    """nq x k row-major results (reference search.hpp:77-84); rows sorted by
    (score desc, id asc)."""

    scores: np.ndarray
    ids: np.ndarray

    @property
    def k(self) -> int:
        return self.scores.shape[1]

    @property
    def nq(self) -> int:
        return self.scores.shape[0]


class TurboQuantIndex:
    def __init__(self, dim: int, bits: int = 4):
        h = P()
        check(lib.vc_index_create(S(dim), S(bits), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            check(lib.vc_index_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference accessors ------------------------------------------------
    def dim(self) -> int:
        return int(lib.vc_index_dim(self._h))

    def size(self) -> int:
        return int(lib.vc_index_size(self._h))

    def bit_width(self) -> int:
        return int(lib.vc_index_bits(self._h))

    def codes_bytes(self) -> int:
        return int(lib.vc_index_codes_bytes(self._h))

    def index_bytes(self) -> int:
        return int(lib.vc_index_index_bytes(self._h))

    @property
    def prepared(self) -> bool:
        return bool(lib.vc_index_is_prepared(self._h))

    @property
    def calibrated(self) -> bool:
        return bool(lib.vc_index_is_calibrated(self._h))

    # ---- build ----------------------------------------------------------------
    def calibrate(self, sample):
        """TQ+ calibration from a sample (fit_calibration semantics)."""
        if _is_torch(sample):
            t = sample.contiguous().float()
            check(lib.vc_index_calibrate_device(self._h, P(t.data_ptr()), S(t.numel() // self.dim())))
            return
        a = _f32(sample, self.dim())
        check(lib.vc_index_calibrate(self._h, _ptr(a), S(a.shape[0])))

    def add(self, vectors, stream=None):
        """Append rows (numpy / host, or a CUDA tensor read in order on
        `stream`, default torch's current stream; no device-wide sync)."""
        if _is_torch(vectors) and vectors.is_cuda:
            t = _cuda_rows(vectors, self.dim(), "vectors")
            check(lib.vc_index_add_device(self._h, P(t.data_ptr()), S(t.numel() // self.dim()),
                                          _launch_stream(stream, t)))
            return
        a = _f32(vectors, self.dim())
        check(lib.vc_index_add(self._h, _ptr(a), S(a.shape[0])))

    def prepare(self):
        check(lib.vc_index_prepare(self._h))

    # ---- search ---------------------------------------------------------------
    def search(self, queries, k: int) -> SearchResults:
        q = _f32(queries, self.dim())
        nq = q.shape[0]
        kk = min(int(k), self.size())
        scores = np.zeros((nq, max(kk, 0)), np.float32)
        ids = np.zeros((nq, max(kk, 0)), np.uint64)
        kout = S()
        check(lib.vc_index_search(self._h, _ptr(q), S(nq), S(k), _ptr(scores), _ptr(ids), ctypes.byref(kout)))
        return SearchResults(scores[:, : kout.value], ids[:, : kout.value])

    def search_device(self, queries, k: int, scores=None, ids=None, stream=None):
        """Device-resident search: queries is a torch CUDA tensor (nq x dim).
        Returns torch tensors (scores f32, ids int64 holding the u64 ids),
        written in order on `stream` (default: torch's current stream)."""
        import torch

        q = _cuda_rows(queries, self.dim(), "queries")
        nq = q.numel() // self.dim()
        kk = min(int(k), self.size())
        scores = _cuda_out(scores, nq, kk, torch.float32, "scores", q.device)
        ids = _cuda_out(ids, nq, kk, torch.int64, "ids", q.device)
        kout = S()
        st = _launch_stream(stream, q)
        check(lib.vc_index_search_device(self._h, P(q.data_ptr()), S(nq), S(k), P(scores.data_ptr()),
                                         P(ids.data_ptr()), ctypes.byref(kout), st))
        return scores, ids

    def search_prefix_device(self, queries, k: int, n_prefix: int, stream=None):
        """search_device over the first n_prefix vectors only: what a search
        issued while the index held n_prefix vectors returned."""
        import torch

        q = _cuda_rows(queries, self.dim(), "queries")
        nq = q.numel() // self.dim()
        kk = min(int(k), int(n_prefix))
        scores = torch.empty((nq, kk), dtype=torch.float32, device=q.device)
        ids = torch.empty((nq, kk), dtype=torch.int64, device=q.device)
        kout = S()
        check(lib.vc_index_search_prefix_device(self._h, P(q.data_ptr()), S(nq), S(k), S(n_prefix),
                                                P(scores.data_ptr()), P(ids.data_ptr()), ctypes.byref(kout),
                                                _launch_stream(stream, q)))
        return scores, ids

    # ---- exports in the reference formats ---------------------------------------
    def export_packed(self) -> np.ndarray:
        out = np.zeros((self.size(), self.bit_width() * self.dim() // 8), np.uint8)
        check(lib.vc_index_export_packed(self._h, _ptr(out)))
        return out

    def export_flat(self) -> np.ndarray:
        G = self.dim() // (8 // self.bit_width())
        out = np.zeros((self.size(), G), np.uint8)
        check(lib.vc_index_export_flat(self._h, _ptr(out)))
        return out

    def export_perm0(self) -> np.ndarray:
        G = self.dim() // (8 // self.bit_width())
        nb = (self.size() + 191) // 192
        out = np.zeros(nb * G * 192, np.uint8)
        check(lib.vc_index_export_perm0(self._h, _ptr(out)))
        return out

    def export_gpu32(self) -> np.ndarray:
        """The reference CUDA index's code words: u32 [ceil(n/32)][ceil(G/4)][32]
        (src/gpu/kernels.cu:111-133), flattened."""
        out = np.zeros(int(lib.vc_gpu32_words(S(self.size()), S(self.dim()), S(self.bit_width()))), np.uint32)
        check(lib.vc_index_export_gpu32(self._h, _ptr(out)))
        return out

    def export_seq(self) -> np.ndarray:
        """repack_seq bytes: sequential 192x1 blocks."""
        G = self.dim() // (8 // self.bit_width())
        nb = (self.size() + 191) // 192
        out = np.zeros(nb * G * 192, np.uint8)
        check(lib.vc_index_export_seq(self._h, _ptr(out)))
        return out

    def export_alpha(self) -> np.ndarray:
        out = np.zeros(self.size(), np.float32)
        check(lib.vc_index_export_alpha(self._h, _ptr(out)))
        return out

    def calibration(self):
        sh = np.zeros(self.dim(), np.float32)
        sc = np.zeros(self.dim(), np.float32)
        check(lib.vc_index_get_calibration(self._h, _ptr(sh), _ptr(sc)))
        return sh, sc

    def prepare_queries(self, queries):
        """prepare_queries state: dict(luts, lut_scale, lut_bias, bias_corr)."""
        q = _f32(queries, self.dim())
        nq = q.shape[0]
        G = self.dim() // (8 // self.bit_width())
        luts = np.zeros((nq, G * 32), np.uint8)
        ls = np.zeros(nq, np.float32)
        lb = np.zeros(nq, np.float32)
        bc = np.zeros(nq, np.float32)
        check(lib.vc_index_prepare_queries(self._h, _ptr(q), S(nq), _ptr(luts), _ptr(ls), _ptr(lb), _ptr(bc)))
        return dict(luts=luts, lut_scale=ls, lut_bias=lb, bias_corr=bc)


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


# ---- stateless functions ---------------------------------------------------------
def rotation_tables(dim: int):
    perm = np.zeros(dim, np.uint32)
    signs = np.zeros(dim, np.float32)
    isb = ctypes.c_float()
    blk = S()
    check(lib.vc_rotation_tables(S(dim), _ptr(perm), _ptr(signs), ctypes.byref(isb), ctypes.byref(blk)))
    return perm, signs, np.float32(isb.value), int(blk.value)


def codebook(bits: int, dim: int):
    b = np.zeros(15, np.float32)
    c = np.zeros(16, np.float32)
    check(lib.vc_codebook(S(bits), S(dim), _ptr(b), _ptr(c)))
    L = 1 << bits
    return b[: L - 1].copy(), c[:L].copy()


def encode(vectors, bits: int, shift=None, scale=None):
    """encode(): bit-plane codes (n x bits*dim/8) and alpha (n)."""
    x = _f32(vectors)
    n, dim = x.shape
    packed = np.zeros((n, bits * dim // 8), np.uint8)
    alpha = np.zeros(n, np.float32)
    sh = None if shift is None else _f32(shift)
    sc = None if scale is None else _f32(scale)
    check(lib.vc_encode(_ptr(x), S(n), S(dim), S(bits), _ptr(sh), _ptr(sc), _ptr(packed), _ptr(alpha)))
    return packed, alpha


def rotate(vectors):
    """rotate_batch_into: rows scaled by 1/||row|| and rotated (n x dim)."""
    x = _f32(vectors)
    n, dim = x.shape
    out = np.zeros((n, dim), np.float32)
    check(lib.vc_rotate(_ptr(x), S(n), S(dim), _ptr(out)))
    return out


def fit_calibration(sample, bits: int):
    x = _f32(sample)
    n, dim = x.shape
    shift = np.zeros(dim, np.float32)
    scale = np.zeros(dim, np.float32)
    check(lib.vc_fit_calibration(_ptr(x), S(n), S(dim), S(bits), _ptr(shift), _ptr(scale)))
    return shift, scale


def search_perm0(queries, k: int, bits: int, perm0, alpha, id_map=None, shift=None, scale=None) -> SearchResults:
    """score_prepared over PERM0 blocked codes (optionally with an id map)."""
    q = _f32(queries)
    nq, dim = q.shape
    alpha = _f32(alpha)
    n = alpha.size
    perm0 = np.ascontiguousarray(perm0, np.uint8)
    kk = min(int(k), n)
    scores = np.zeros((nq, kk), np.float32)
    ids = np.zeros((nq, kk), np.uint64)
    im = None if id_map is None else np.ascontiguousarray(id_map, np.uint64)
    sh = None if shift is None else _f32(shift)
    sc = None if scale is None else _f32(scale)
    kout = S()
    check(lib.vc_search_perm0(_ptr(q), S(nq), S(dim), S(bits), S(k), _ptr(perm0), _ptr(alpha), S(n), _ptr(im),
                              _ptr(sh), _ptr(sc), _ptr(scores), _ptr(ids), ctypes.byref(kout)))
    return SearchResults(scores[:, : kout.value], ids[:, : kout.value])


def score_perm0(queries, k: int, bits: int, perm0, alpha, id_map=None, shift=None, scale=None,
                query_indices=None, residual_norms=None, residual_scales=None) -> SearchResults:
    """score_prepared over PERM0 blocked codes with every argument of the
    reference's AVX2 path: id map, query subset (rows in its order) and the
    residual post-correction (residual_norms per vector, residual_scales per
    query)."""
    q = _f32(queries)
    nq, dim = q.shape
    alpha = _f32(alpha)
    n = alpha.size
    perm0 = np.ascontiguousarray(perm0, np.uint8)
    qi = None if query_indices is None else np.ascontiguousarray(query_indices, np.uint64)
    nq_out = nq if qi is None else qi.size
    kk = min(int(k), n)
    scores = np.zeros((nq_out, kk), np.float32)
    ids = np.zeros((nq_out, kk), np.uint64)
    cv = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
    kout = S()
    check(lib.vc_score_perm0(_ptr(q), S(nq), S(dim), S(bits), S(k), _ptr(perm0), _ptr(alpha), S(n),
                             _ptr(cv(id_map, np.uint64)), _ptr(cv(shift, np.float32)), _ptr(cv(scale, np.float32)),
                             _ptr(qi), S(0 if qi is None else qi.size), _ptr(cv(residual_norms, np.float32)),
                             _ptr(cv(residual_scales, np.float32)), _ptr(scores), _ptr(ids), ctypes.byref(kout)))
    return SearchResults(scores[:, : kout.value], ids[:, : kout.value])


def gpu32_to_perm0(codes, n: int, dim: int, bits: int) -> np.ndarray:
    """Codes in the reference CUDA index's layout (u32 words
    [ceil(n/32)][ceil(G/4)][32]) as PERM0 blocks, the input of search_perm0 and
    score_perm0_into."""
    words = int(lib.vc_gpu32_words(S(n), S(dim), S(bits)))
    codes = np.ascontiguousarray(codes, np.uint32).reshape(-1)
    if words == 0 or codes.size != words:
        raise ValueError(f"expected {words} code words for n={n}, dim={dim}, bits={bits}, got {codes.size}")
    G = dim // (8 // bits)
    out = np.zeros((n + 191) // 192 * G * 192, np.uint8)
    check(lib.vc_gpu32_to_perm0(_ptr(codes), S(n), S(dim), S(bits), _ptr(out)))
    return out


def empty_heaps(nq: int, k: int):
    """Fresh per-query heaps in the reference's layout: (heap_s [nq, k] f32,
    heap_i [nq, k] u64, heap_sz [nq] u64, heap_min [nq] f32 = -inf,
    heap_mi [nq] u64)."""
    return (np.zeros((nq, k), np.float32), np.zeros((nq, k), np.uint64), np.zeros(nq, np.uint64),
            np.full(nq, -np.inf, np.float32), np.zeros(nq, np.uint64))


def score_perm0_into(queries, k: int, bits: int, perm0, alpha, query_indices, heaps, id_map=None, shift=None,
                     scale=None, score_offsets=None):
    """score_prepared_into (reference search.cpp:1156-1267) over one bucket of
    PERM0 blocks + alpha: the listed queries' heaps (see empty_heaps) are
    updated in place and returned."""
    q = _f32(queries)
    nq, dim = q.shape
    alpha = _f32(alpha)
    n = alpha.size
    perm0 = np.ascontiguousarray(perm0, np.uint8)
    qi = np.ascontiguousarray(query_indices, np.uint64)
    hs, hi, hsz, hmin, hmi = heaps
    for a, dt in ((hs, np.float32), (hi, np.uint64), (hsz, np.uint64), (hmin, np.float32), (hmi, np.uint64)):
        if a.dtype != dt or not a.flags.c_contiguous:
            raise ValueError("heaps must be C-contiguous arrays of the empty_heaps() types")
    if hs.shape != (nq, k) or hi.shape != (nq, k) or hsz.shape != (nq,) or hmin.shape != (nq,) or hmi.shape != (nq,):
        raise ValueError("heap shapes must be [nq, k] and [nq]")
    im = None if id_map is None else np.ascontiguousarray(id_map, np.uint64)
    sh = None if shift is None else _f32(shift)
    sc = None if scale is None else _f32(scale)
    off = None if score_offsets is None else _f32(score_offsets)
    check(lib.vc_score_perm0_into(_ptr(q), S(nq), S(dim), S(bits), S(k), _ptr(perm0), _ptr(alpha), S(n), _ptr(im),
                                  _ptr(sh), _ptr(sc), _ptr(qi), S(qi.size), _ptr(hs), _ptr(hi), _ptr(hsz), _ptr(hmin),
                                  _ptr(hmi), _ptr(off)))
    return heaps


class PreparedQueries:
    """prepare_queries (reference search.cpp:762-860) run once on the GPU and
    kept in device memory, reused by score_prepared / score_prepared_into over
    many code ranges as the reference reuses its PreparedQueries
    (include/vectorcache/search/search.hpp:95-112)."""

    def __init__(self, handle):
        self._h = handle
        nq, dim, bits, g = S(), S(), S(), S()
        check(lib.vc_prepared_shape(self._h, ctypes.byref(nq), ctypes.byref(dim), ctypes.byref(bits), ctypes.byref(g)))
        self.nq, self.dim, self.bits, self.n_byte_groups = nq.value, dim.value, bits.value, g.value

    def close(self):
        if self._h:
            lib.vc_prepared_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self) -> dict:
        """luts [nq, G*32] u8, lut_scale, lut_bias (without bias_corr), bias_corr."""
        g = self.n_byte_groups
        out = {"luts": np.zeros((self.nq, g * 32), np.uint8), "lut_scale": np.zeros(self.nq, np.float32),
               "lut_bias": np.zeros(self.nq, np.float32), "bias_corr": np.zeros(self.nq, np.float32)}
        check(lib.vc_prepared_export(self._h, _ptr(out["luts"]), _ptr(out["lut_scale"]), _ptr(out["lut_bias"]),
                                     _ptr(out["bias_corr"])))
        return out

    def query_at(self, qi: int) -> "PreparedQueries":
        """prepared_query_at (reference search.cpp:1045-1068)."""
        h = P()
        check(lib.vc_prepared_query_at(self._h, S(qi), ctypes.byref(h)))
        return PreparedQueries(h)


def prepare_queries(queries, bits: int, shift=None, scale=None) -> PreparedQueries:
    """Host query rows (nq x dim) -> PreparedQueries on the GPU."""
    q = _f32(queries)
    nq, dim = q.shape
    sh = None if shift is None else _f32(shift)
    sc = None if scale is None else _f32(scale)
    h = P()
    check(lib.vc_prepare_queries(_ptr(q), S(nq), S(dim), S(bits), _ptr(sh), _ptr(sc), ctypes.byref(h)))
    return PreparedQueries(h)


def score_prepared(prep: PreparedQueries, k: int, perm0, alpha, id_map=None, query_indices=None, residual_norms=None,
                   residual_scales=None) -> SearchResults:
    """score_prepared (reference search.cpp:1072-1154) with prepared queries:
    vc_score_perm0's results without rebuilding the LUTs."""
    alpha = _f32(alpha)
    n = alpha.size
    perm0 = np.ascontiguousarray(perm0, np.uint8)
    im = None if id_map is None else np.ascontiguousarray(id_map, np.uint64)
    qi = None if query_indices is None else np.ascontiguousarray(query_indices, np.uint64)
    rn = None if residual_norms is None else _f32(residual_norms)
    rs = None if residual_scales is None else _f32(residual_scales)
    nq_out = prep.nq if qi is None else qi.size
    ke = min(k, n)
    scores = np.zeros((nq_out, max(ke, 0)), np.float32)
    ids = np.zeros((nq_out, max(ke, 0)), np.uint64)
    ko = S()
    check(lib.vc_score_prepared(prep._h, S(k), _ptr(perm0), _ptr(alpha), S(n), _ptr(im), _ptr(qi),
                                S(0 if qi is None else qi.size), _ptr(rn), _ptr(rs), _ptr(scores), _ptr(ids),
                                ctypes.byref(ko)))
    return SearchResults(scores[:, :ko.value], ids[:, :ko.value])


def score_prepared_into(prep: PreparedQueries, k: int, perm0, alpha, query_indices, heaps, id_map=None,
                        score_offsets=None):
    """score_prepared_into (reference search.cpp:1156-1267) with prepared
    queries; the listed queries' heaps (see empty_heaps) are updated in place."""
    nq = prep.nq
    alpha = _f32(alpha)
    n = alpha.size
    perm0 = np.ascontiguousarray(perm0, np.uint8)
    qi = np.ascontiguousarray(query_indices, np.uint64)
    hs, hi, hsz, hmin, hmi = heaps
    for a, dt in ((hs, np.float32), (hi, np.uint64), (hsz, np.uint64), (hmin, np.float32), (hmi, np.uint64)):
        if a.dtype != dt or not a.flags.c_contiguous:
            raise ValueError("heaps must be C-contiguous arrays of the empty_heaps() types")
    if hs.shape != (nq, k) or hi.shape != (nq, k) or hsz.shape != (nq,) or hmin.shape != (nq,) or hmi.shape != (nq,):
        raise ValueError("heap shapes must be [nq, k] and [nq]")
    im = None if id_map is None else np.ascontiguousarray(id_map, np.uint64)
    off = None if score_offsets is None else _f32(score_offsets)
    check(lib.vc_score_prepared_into(prep._h, S(k), _ptr(perm0), _ptr(alpha), S(n), _ptr(im), _ptr(qi), S(qi.size),
                                     _ptr(hs), _ptr(hi), _ptr(hsz), _ptr(hmin), _ptr(hmi), _ptr(off)))
    return heaps


def _repack(fn, packed, n: int, dim: int, bits: int, out_bytes: int) -> np.ndarray:
    packed = np.ascontiguousarray(packed, np.uint8)
    if packed.size != n * bits * dim // 8:
        raise ValueError("packed must hold n x bits*dim/8 bytes")
    out = np.zeros(out_bytes, np.uint8)
    check(fn(_ptr(packed), S(n), S(dim), S(bits), _ptr(out)))
    return out


def codes_flat_from_packed(packed, n: int, dim: int, bits: int) -> np.ndarray:
    """codes_flat_from_packed (reference pack.cpp:98-131,352-356) on the GPU: [n, G] bytes."""
    g = dim // (8 // bits)
    return _repack(lib.vc_codes_flat_from_packed, packed, n, dim, bits, n * g).reshape(n, g)


def repack_perm0(packed, n: int, dim: int, bits: int) -> np.ndarray:
    """repack_perm0 (reference pack.cpp:157-186,388-401) on the GPU."""
    g = dim // (8 // bits)
    return _repack(lib.vc_repack_perm0, packed, n, dim, bits, (n + 191) // 192 * g * 192)


def repack_seq(packed, n: int, dim: int, bits: int) -> np.ndarray:
    """repack_seq (reference pack.cpp:133-154,358-386) on the GPU."""
    g = dim // (8 // bits)
    return _repack(lib.vc_repack_seq, packed, n, dim, bits, (n + 191) // 192 * g * 192)


def heaps_to_results(heaps) -> SearchResults:
    """heaps_to_search_results (reference search.cpp:1269-1291) on the GPU."""
    hs, hi, hsz = heaps[0], heaps[1], heaps[2]
    nq, k = hs.shape
    scores = np.zeros((nq, k), np.float32)
    ids = np.zeros((nq, k), np.uint64)
    check(lib.vc_heaps_to_results(S(nq), S(k), _ptr(np.ascontiguousarray(hs, np.float32)),
                                  _ptr(np.ascontiguousarray(hi, np.uint64)), _ptr(np.ascontiguousarray(hsz, np.uint64)),
                                  _ptr(scores), _ptr(ids)))
    return SearchResults(scores, ids)


def float_rescore_topk(results: SearchResults, queries, corpus, counts=None) -> SearchResults:
    """float_rescore_topk (reference float_rescore.cpp:80-164) on the GPU:
    re-scored, re-ordered copies of the rows. With torch CUDA tensors
    (scores f32, ids int64, queries, corpus) it runs in place on the device,
    ordered on torch's current stream."""
    if _is_torch(results.scores) and results.scores.is_cuda:
        import torch

        sc, ii = results.scores, results.ids
        nq, k = sc.shape
        q = queries.contiguous().float()
        cor = corpus.contiguous().float()
        cnt = None if counts is None else counts.contiguous().to(torch.int64)
        st = torch.cuda.current_stream().cuda_stream
        check(lib.vc_float_rescore_device(S(nq), S(k), P(sc.data_ptr()), P(ii.data_ptr()),
                                          None if cnt is None else P(cnt.data_ptr()), P(q.data_ptr()),
                                          P(cor.data_ptr()), S(cor.shape[0]), S(q.shape[1]), P(st)))
        return results
    sc = np.ascontiguousarray(results.scores, np.float32).copy()
    ii = np.ascontiguousarray(results.ids, np.uint64).copy()
    nq, k = sc.shape
    q = _f32(queries)
    cor = _f32(corpus)
    cnt = None if counts is None else np.ascontiguousarray(counts, np.uint64)
    check(lib.vc_float_rescore(S(nq), S(k), _ptr(sc), _ptr(ii), _ptr(cnt), _ptr(q), _ptr(cor), S(cor.shape[0]),
                               S(q.shape[1])))
    return SearchResults(sc, ii)


def merge_search_results(dst: SearchResults, src: SearchResults) -> SearchResults:
    """merge_search_results (reference search.cpp:1551-1600) on the GPU."""
    if src.nq == 0 or src.k == 0:
        return dst
    if dst.nq == 0 or dst.k == 0:
        return SearchResults(src.scores.copy(), src.ids.copy())
    if dst.nq != src.nq:
        raise ValueError("merge_search_results: nq mismatch")
    ds = np.ascontiguousarray(dst.scores, np.float32).copy()
    di = np.ascontiguousarray(dst.ids, np.uint64).copy()
    ss = np.ascontiguousarray(src.scores, np.float32)
    si = np.ascontiguousarray(src.ids, np.uint64)
    check(lib.vc_merge_results(S(dst.nq), S(dst.k), _ptr(ds), _ptr(di), S(src.k), _ptr(ss), _ptr(si)))
    return SearchResults(ds, di)


def set_scan_tile(tile: int):
    check(lib.vc_set_scan_tile(S(tile)))


def scan_tile_for(dim: int, bits: int, nq: int) -> int:
    return int(lib.vc_scan_tile_for(S(dim), S(bits), S(nq)))


def set_debug(flags: int):
    """Test hooks: bit 0 forces the exact sequential alpha path in encode;
    bit 2 selects the tile encode kernel for indexes created afterwards;
    bit 1 selects the nibble-table scan kernel instead of the byte-pair one;
    bit 3 swaps the byte-pair kernel's vectors-per-thread variant (1 <-> 2);
    bit 4 forces the sequential bias_corr chain in query preparation;
    bit 5 makes landmark searches rotate queries in a separate kernel;
    bit 7 / bit 8 make bucketed searches take the bucket-major tile scan /
    the query-major scan whatever the batch (k <= 64 for the tile scan);
    bits 16-31 (when not zero) force the number of vector ranges per query
    tile (at most 512)."""
    check(lib.vc_set_debug(ctypes.c_uint(flags)))


def set_profiling(on: bool):
    check(lib.vc_set_profiling(1 if on else 0))


def last_scan_stats():
    """(ms, tile, n_ranges) of the last profiled scan kernel (HIP events)."""
    ms = ctypes.c_double()
    tile = S()
    ranges = S()
    check(lib.vc_last_scan_stats(ctypes.byref(ms), ctypes.byref(tile), ctypes.byref(ranges)))
    return ms.value, int(tile.value), int(ranges.value)


version = _capi.version
