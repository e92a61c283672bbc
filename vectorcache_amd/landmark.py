"""The landmark index: the reference's LandmarkIndex (include/vectorcache/index.hpp:16-68,
src/index.cpp:120-289), which its TurboQuantIndex aliases, on the AVX2 path.

add() encodes alpha * y-hat of each vector, projects it onto an orthonormal
landmark basis and keeps the norm of the residual; prepare() encodes the
landmark rows (scan_dim = n_landmarks rounded up to 8) for the LUT scan;
search() rotates the queries, projects them onto the 4-bit basis, scans and
applies the residual post-correction. Every step runs on the GPU through
libvectorcache_hip.so; results equal the reference's bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._capi import P, S, check, lib
from .index import SearchResults, _cuda_out, _cuda_rows, _f32, _is_torch, _launch_stream, _ptr

DEFAULT_LANDMARKS = 1024  # reference kDefaultLandmarks (include/vectorcache/constants.hpp:28)


class LandmarkIndex:
    def __init__(self, dim: int, scan_bits: int, n_landmarks: int = DEFAULT_LANDMARKS, landmark_seed: int = 0,
                 rerank_bits: int = 0, planes: bool = False):
        """planes=True: the reference's behaviour on an AVX-512 VNNI host (sign
        planes from 32768 prepared vectors, the staged sign-plane search)."""
        h = P()
        check(lib.vc_lm_index_create(S(dim), S(scan_bits), S(n_landmarks), ctypes.c_uint64(landmark_seed),
                                     S(rerank_bits), ctypes.byref(h)))
        self._h = h
        if planes:
            self.set_planes(True)

    def set_planes(self, on: bool):
        check(lib.vc_lm_index_set_planes(self._h, ctypes.c_int(int(on))))

    def planes_active(self) -> bool:
        """Whether the index holds sign planes (searches may take the staged branch)."""
        return bool(lib.vc_lm_index_planes_active(self._h))

    def close(self):
        if getattr(self, "_h", None):
            check(lib.vc_lm_index_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference accessors ------------------------------------------------
    def dim(self) -> int:
        return int(lib.vc_lm_index_dim(self._h))

    def bit_width(self) -> int:
        return int(lib.vc_lm_index_scan_bits(self._h))

    def scan_bits(self) -> int:
        return int(lib.vc_lm_index_scan_bits(self._h))

    def rerank_bits(self) -> int:
        return int(lib.vc_lm_index_rerank_bits(self._h))

    def n_landmarks(self) -> int:
        return int(lib.vc_lm_index_n_landmarks(self._h))

    def scan_dim(self) -> int:
        return int(lib.vc_lm_index_scan_dim(self._h))

    def size(self) -> int:
        return int(lib.vc_lm_index_size(self._h))

    def codes_bytes(self) -> int:
        return int(lib.vc_lm_index_codes_bytes(self._h))

    def index_bytes(self) -> int:
        return int(lib.vc_lm_index_index_bytes(self._h))

    @property
    def prepared(self) -> bool:
        return bool(lib.vc_lm_index_is_prepared(self._h))

    # ---- build ----------------------------------------------------------------
    def add(self, vectors, stream=None):
        if _is_torch(vectors) and vectors.is_cuda:
            t = vectors.contiguous().float()
            if t.numel() % self.dim() != 0:
                raise ValueError("vectors length must be multiple of dim")
            check(lib.vc_lm_index_add_device(self._h, P(t.data_ptr()), S(t.numel() // self.dim()),
                                             _launch_stream(stream, t)))
            return
        a = np.ascontiguousarray(vectors, dtype=np.float32)
        if a.size % self.dim() != 0:
            raise ValueError("vectors length must be multiple of dim")
        a = a.reshape(-1, self.dim())
        check(lib.vc_lm_index_add(self._h, _ptr(a), S(a.shape[0])))

    def prepare(self):
        check(lib.vc_lm_index_prepare(self._h))

    # ---- search ---------------------------------------------------------------
    def search(self, queries, k: int, float_corpus=None) -> SearchResults:
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.size % self.dim() != 0:
            raise ValueError("queries length must be multiple of dim")
        q = q.reshape(-1, self.dim())
        fc = None
        if float_corpus is not None:
            fc = np.ascontiguousarray(float_corpus, dtype=np.float32)
            if fc.size < self.size() * self.dim():
                raise ValueError("float_corpus must have at least size() * dim rows")
        nq = q.shape[0]
        kk = min(int(k), self.size())
        scores = np.zeros((nq, kk), np.float32)
        ids = np.zeros((nq, kk), np.uint64)
        kout = S()
        check(lib.vc_lm_index_search(self._h, _ptr(q), S(nq), S(k), _ptr(fc), _ptr(scores), _ptr(ids),
                                     ctypes.byref(kout)))
        return SearchResults(scores[:, : kout.value], ids[:, : kout.value])

    def search_device(self, queries, k: int, scores=None, ids=None, float_corpus=None, stream=None):
        """Device-resident search: queries (and the optional float corpus) are
        torch CUDA tensors. Returns torch tensors (scores f32, ids int64 holding
        the u64 ids), written in order on `stream` (default: torch's current
        stream)."""
        import torch

        q = _cuda_rows(queries, self.dim(), "queries")
        nq = q.numel() // self.dim()
        kk = min(int(k), self.size())
        scores = _cuda_out(scores, nq, kk, torch.float32, "scores", q.device)
        ids = _cuda_out(ids, nq, kk, torch.int64, "ids", q.device)
        fc = None
        if float_corpus is not None:
            fc = _cuda_rows(float_corpus, self.dim(), "float_corpus")
            if fc.numel() < self.size() * self.dim():
                raise ValueError("float_corpus must have at least size() * dim rows")
        kout = S()
        st = _launch_stream(stream, q, fc)
        check(lib.vc_lm_index_search_device(self._h, P(q.data_ptr()), S(nq), S(k),
                                            P(fc.data_ptr()) if fc is not None else None, P(scores.data_ptr()),
                                            P(ids.data_ptr()), ctypes.byref(kout), st))
        return scores, ids

    # ---- inspection -------------------------------------------------------------
    def query_rows(self, queries):
        """The query side of search(): landmark rows (nq x scan_dim) and the
        residual scales |q_perp| (nq)."""
        q = _f32(queries, self.dim())
        nq = q.shape[0]
        ql = np.zeros((nq, self.scan_dim()), np.float32)
        qp = np.zeros(nq, np.float32)
        check(lib.vc_lm_index_query_rows(self._h, _ptr(q), S(nq), _ptr(ql), _ptr(qp)))
        return ql, qp

    def residual_norms(self) -> np.ndarray:
        rn = np.zeros(self.size(), np.float32)
        check(lib.vc_lm_index_export(self._h, _ptr(rn), None, None))
        return rn

    def export_codes(self):
        """(alpha, flat codes n x scan_dim / (8 / scan_bits)) of the prepared rows."""
        n = self.size()
        g = self.scan_dim() // (8 // self.scan_bits())
        alpha = np.zeros(n, np.float32)
        flat = np.zeros((n, g), np.uint8)
        check(lib.vc_lm_index_export(self._h, None, _ptr(alpha), _ptr(flat)))
        return alpha, flat


def make_ortho_landmarks(dim: int, m: int, seed: int = 0) -> np.ndarray:
    """make_ortho_landmarks (reference src/transform/landmarks.cpp:180-198):
    returns L as m x dim (row j = landmark column j)."""
    L = np.zeros((m, dim), np.float32)
    check(lib.vc_landmarks_make(S(dim), S(m), ctypes.c_uint64(seed), _ptr(L)))
    return L


def quantize_landmarks(L) -> tuple[np.ndarray, np.ndarray]:
    """quantize_landmarks (reference src/transform/landmarks.cpp:236-269) of L
    given as m x dim: (codes m x pad16(dim) / 2, scales m)."""
    L = np.ascontiguousarray(L, np.float32)
    m, dim = L.shape
    codes = np.zeros((m, ((dim + 15) // 16 * 16) // 2), np.uint8)
    scales = np.zeros(m, np.float32)
    check(lib.vc_landmarks_quantize(_ptr(L), S(dim), S(m), _ptr(codes), _ptr(scales)))
    return codes, scales
