"""BucketedTurboQuantIndex: the README's bucketed index (reference README.md:56,
114-120: BucketParams, calibrate/add/prepare/search) on the GPU.

Rows are L2-normalized and grouped into buckets by a coarse quantizer; each
bucket's rows are encoded with the flat index's encode (as residuals against
the bucket's encode centroid by default); a query opens its best buckets by
routing score and every opened bucket is scored as score_prepared_into scores
one bucket (reference src/search/search.cpp:1156-1267), with the offset
<q_hat, c> added after alpha. The number of buckets a query opens is the
fixed n_probe, a per-query count, or the learned probe model's choice.
Every call goes through the C ABI of libvectorcache_hip.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import P, S, check, lib
from .index import SearchResults, _cuda_out, _cuda_rows, _f32, _is_torch, _launch_stream, _ptr

PROBE_FEATURES = 96
PROBE_MAX = 256


@dataclass
class BucketParams:
    """vc_bucket_params (the README's BucketParams fields n_probe,
    max_bucket_size, energy_soft, expected_n, plus the build's own)."""

    n_probe: int = 64
    max_bucket_size: int = 800
    energy_soft: int = 400
    expected_n: int = 0
    kmeans_iters: int = 8
    kmeans_sample: int = 64
    seed: int = 4001
    residual: bool = True

    def _c(self):
        p = _capi.BucketParams()
        lib.vc_bucket_params_default(ctypes.byref(p))
        p.n_probe, p.max_bucket_size, p.energy_soft = self.n_probe, self.max_bucket_size, self.energy_soft
        p.expected_n, p.kmeans_iters, p.kmeans_sample = self.expected_n, self.kmeans_iters, self.kmeans_sample
        p.seed, p.residual = self.seed, 1 if self.residual else 0
        return p


class BucketedTurboQuantIndex:
    def __init__(self, dim: int, bits: int = 4, params: BucketParams | None = None):
        self.params = params or BucketParams()
        h = P()
        cp = self.params._c()
        check(lib.vc_bucketed_create(S(dim), S(bits), ctypes.byref(cp), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            check(lib.vc_bucketed_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- accessors ----------------------------------------------------------------
    def dim(self) -> int:
        return int(lib.vc_bucketed_dim(self._h))

    def size(self) -> int:
        return int(lib.vc_bucketed_size(self._h))

    def bit_width(self) -> int:
        return int(lib.vc_bucketed_bits(self._h))

    def n_buckets(self) -> int:
        return int(lib.vc_bucketed_n_buckets(self._h))

    @property
    def prepared(self) -> bool:
        return bool(lib.vc_bucketed_is_prepared(self._h))

    @property
    def calibrated(self) -> bool:
        return bool(lib.vc_bucketed_is_calibrated(self._h))

    # ---- build --------------------------------------------------------------------
    def calibrate(self, sample):
        if _is_torch(sample) and sample.is_cuda:
            t = _cuda_rows(sample, self.dim(), "sample")
            check(lib.vc_bucketed_calibrate_device(self._h, P(t.data_ptr()), S(t.numel() // self.dim()),
                                                   _launch_stream(None)))
            return
        a = _f32(sample, self.dim())
        check(lib.vc_bucketed_calibrate(self._h, _ptr(a), S(a.shape[0])))

    def add(self, vectors):
        if _is_torch(vectors) and vectors.is_cuda:
            t = _cuda_rows(vectors, self.dim(), "vectors")
            check(lib.vc_bucketed_add_device(self._h, P(t.data_ptr()), S(t.numel() // self.dim()),
                                             _launch_stream(None)))
            return
        a = _f32(vectors, self.dim())
        check(lib.vc_bucketed_add(self._h, _ptr(a), S(a.shape[0])))

    def prepare(self):
        check(lib.vc_bucketed_prepare(self._h))

    # ---- search -------------------------------------------------------------------
    def search(self, queries, k: int) -> SearchResults:
        q = _f32(queries, self.dim())
        nq = q.shape[0]
        kk = min(int(k), self.size())
        scores = np.zeros((nq, kk), np.float32)
        ids = np.zeros((nq, kk), np.uint64)
        kout = S()
        check(lib.vc_bucketed_search(self._h, _ptr(q), S(nq), S(k), _ptr(scores), _ptr(ids), ctypes.byref(kout)))
        return SearchResults(scores[:, : kout.value], ids[:, : kout.value])

    def search_device(self, queries, k: int, n_probe: int = 0, probe_counts=None, scores=None, ids=None,
                      counts=None, stream=None):
        """Device-resident search (torch CUDA tensors). n_probe > 0 opens that
        many buckets per query (or probe_counts[q], an int32 tensor, capped by
        it); n_probe = 0 uses the learned model when enabled, else the
        params' n_probe. Returns (scores f32, ids int64, counts int64 or None)."""
        import torch

        q = _cuda_rows(queries, self.dim(), "queries")
        nq = q.numel() // self.dim()
        kk = min(int(k), self.size())
        scores = _cuda_out(scores, nq, kk, torch.float32, "scores", q.device)
        ids = _cuda_out(ids, nq, kk, torch.int64, "ids", q.device)
        if counts is not None:
            counts = _cuda_out(counts, nq, 1, torch.int64, "counts", q.device)
        pc = None
        if probe_counts is not None:
            pc = probe_counts
            if not (_is_torch(pc) and pc.is_cuda and pc.dtype == torch.int32 and pc.is_contiguous()
                    and pc.numel() >= nq):
                raise ValueError("probe_counts must be a contiguous int32 CUDA tensor of nq entries")
        kout = S()
        st = _launch_stream(stream, q)
        check(lib.vc_bucketed_search_device(self._h, P(q.data_ptr()), S(nq), S(k), S(n_probe),
                                            P(pc.data_ptr()) if pc is not None else None, P(scores.data_ptr()),
                                            P(ids.data_ptr()), P(counts.data_ptr()) if counts is not None else None,
                                            ctypes.byref(kout), st))
        return scores, ids, counts

    def search_exhaustive_device(self, queries, k: int, stream=None):
        """residual = False only: the plain flat scan over every code (the
        judge of the probe budget's recall). Returns (scores f32, ids int64)."""
        import torch

        q = _cuda_rows(queries, self.dim(), "queries")
        nq = q.numel() // self.dim()
        kk = min(int(k), self.size())
        scores = torch.empty((nq, kk), dtype=torch.float32, device=q.device)
        ids = torch.empty((nq, kk), dtype=torch.int64, device=q.device)
        kout = S()
        check(lib.vc_bucketed_search_exhaustive_device(self._h, P(q.data_ptr()), S(nq), S(k), P(scores.data_ptr()),
                                                       P(ids.data_ptr()), ctypes.byref(kout),
                                                       _launch_stream(stream, q)))
        return scores, ids

    def route_device(self, queries, P_: int, stream=None):
        """Each query's P_ best buckets: (bucket ids int32, routing scores f32)."""
        import torch

        q = _cuda_rows(queries, self.dim(), "queries")
        nq = q.numel() // self.dim()
        pb = torch.empty((nq, P_), dtype=torch.int32, device=q.device)
        ps = torch.empty((nq, P_), dtype=torch.float32, device=q.device)
        check(lib.vc_bucketed_route_device(self._h, P(q.data_ptr()), S(nq), S(P_), P(pb.data_ptr()),
                                           P(ps.data_ptr()), _launch_stream(stream, q)))
        return pb, ps

    # ---- built state ----------------------------------------------------------------
    def export(self):
        """dict(bucket_sizes, centroids, pos_ids, flat, alpha, shift, scale):
        positions run bucket by bucket."""
        n, L, D = self.size(), self.n_buckets(), self.dim()
        G = D // (8 // self.bit_width())
        out = dict(bucket_sizes=np.zeros(L, np.uint32), centroids=np.zeros((L, D), np.float32),
                   pos_ids=np.zeros(n, np.uint64), flat=np.zeros((n, G), np.uint8), alpha=np.zeros(n, np.float32),
                   shift=np.zeros(D, np.float32), scale=np.zeros(D, np.float32))
        check(lib.vc_bucketed_export(self._h, *(_ptr(out[k]) for k in ("bucket_sizes", "centroids", "pos_ids", "flat",
                                                                        "alpha", "shift", "scale"))))
        return out

    # ---- learned probe budget -------------------------------------------------------
    def probe_features(self, queries, stream=None):
        import torch

        q = _cuda_rows(queries, self.dim(), "queries")
        nq = q.numel() // self.dim()
        f = torch.empty((nq, PROBE_FEATURES), dtype=torch.float32, device=q.device)
        check(lib.vc_bucketed_probe_features(self._h, P(q.data_ptr()), S(nq), P(f.data_ptr()),
                                             _launch_stream(stream, q)))
        return f

    def probe_init(self, seed: int = 0):
        check(lib.vc_bucketed_probe_init(self._h, ctypes.c_uint64(seed)))

    def probe_weights(self) -> np.ndarray:
        w = np.zeros(int(lib.vc_probe_param_count()), np.float32)
        check(lib.vc_bucketed_probe_get_weights(self._h, _ptr(w)))
        return w

    def set_probe_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != int(lib.vc_probe_param_count()):
            raise ValueError("wrong number of probe parameters")
        check(lib.vc_bucketed_probe_set_weights(self._h, _ptr(w)))

    def probe_train(self, feats, targets, order, lr: float = 1e-3, stream=None):
        """len(order) // 64 Adam steps over minibatches order[64 s .. 64 s + 63]
        (int32 CUDA tensor); returns the minibatch losses (CUDA tensor)."""
        import torch

        n_steps = order.numel() // 64
        losses = torch.empty(max(n_steps, 1), dtype=torch.float32, device=feats.device)
        f = _cuda_rows(feats, PROBE_FEATURES, "feats")
        t = targets.contiguous().float()
        o = order.contiguous().to(torch.int32)
        check(lib.vc_bucketed_probe_train(self._h, P(f.data_ptr()), P(t.data_ptr()), S(t.numel()), P(o.data_ptr()),
                                          S(n_steps), ctypes.c_float(lr), P(losses.data_ptr()),
                                          _launch_stream(stream, f, t, o)))
        return losses[:n_steps]

    def probe_grads(self, feats, targets, order) -> np.ndarray:
        import torch

        f = _cuda_rows(feats, PROBE_FEATURES, "feats")
        t = targets.contiguous().float()
        o = order.contiguous().to(torch.int32)
        g = np.zeros(int(lib.vc_probe_param_count()), np.float32)
        torch.cuda.synchronize()
        check(lib.vc_bucketed_probe_grads(self._h, P(f.data_ptr()), P(t.data_ptr()), S(t.numel()), P(o.data_ptr()),
                                          _ptr(g)))
        return g

    def probe_predict(self, feats, stream=None):
        import torch

        f = _cuda_rows(feats, PROBE_FEATURES, "feats")
        n = f.numel() // PROBE_FEATURES
        y = torch.empty(n, dtype=torch.float32, device=f.device)
        check(lib.vc_bucketed_probe_predict(self._h, P(f.data_ptr()), S(n), P(y.data_ptr()), _launch_stream(stream, f)))
        return y

    def probe_set_table(self, counts, y0: float, y_step: float):
        """Prediction y opens counts[clamp(floor((y - y0) / y_step), 0, len - 1)]
        buckets (empty counts: back to ceil(2^(y + shift)))."""
        c = np.ascontiguousarray(counts, np.int32)
        check(lib.vc_bucketed_probe_set_table(self._h, _ptr(c) if c.size else None, S(c.size), ctypes.c_float(y0),
                                              ctypes.c_float(y_step)))

    def probe_enable(self, on: bool = True, shift: float = 0.0):
        check(lib.vc_bucketed_probe_enable(self._h, 1 if on else 0, ctypes.c_float(shift)))
