"""Request server over a prepared TurboQuantIndex (vc_server_* of the C ABI).

Continuous batching: search requests (query rows, any k up to max_k) and add
batches are submitted from any thread and return a ticket at once; worker
slots, each on its own HIP stream, take the waiting requests in arrival order
into one batch whenever they are free, each batch admitted when its device
memory fits the server's pool; add batches are encoded on a stream of their
own while searches run, and each search reports the index size it saw. No
counterpart in the reference (whose GPU search is not safe to call
concurrently, src/gpu/kernels.cu:602-627); each request's result is that of
TurboQuantIndex::search (src/index.cpp:267-289) over the vectors its batch
scanned.
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import numpy as np

from ._capi import P, RequestInfo, S, ServerParams as _CParams, ServerStats, check, lib
from .index import SearchResults, TurboQuantIndex, _f32, _ptr


@dataclass
class ServerParams:
    slots: int = 4
    max_batch_queries: int = 512
    max_k: int = 100
    hbm_budget_bytes: int = 1 << 30
    reserve_vectors: int = 0  # 0: size() + 1,000,000
    max_vectors: int = 0  # 0: no limit but device memory
    add_chunk: int = 65536

    def to_c(self) -> _CParams:
        return _CParams(*(S(int(getattr(self, f))) for f, _ in _CParams._fields_))


@dataclass
class Served:
    """A finished request: its results (search) and its timing on
    CLOCK_MONOTONIC (time.monotonic())."""

    results: SearchResults | None
    t_submit: float
    t_start: float
    t_done: float
    n_snapshot: int
    k_eff: int
    batch_queries: int
    batch: int
    slot: int


def batch_bytes(index: TurboQuantIndex, n: int, nq: int, k: int) -> int:
    """Pool bytes one batch of nq queries at top-k over n vectors holds."""
    return int(lib.vc_server_batch_bytes(index._h, S(n), S(nq), S(k)))


class Server:
    """Serves `index` (prepared; the index refuses its own add / prepare /
    search / export calls while served). close() finishes every submitted
    request, then returns the index to direct use with the added vectors."""

    def __init__(self, index: TurboQuantIndex, params: ServerParams | None = None):
        self.index = index
        self.params = params or ServerParams()
        self._dim = index.dim()
        h = P()
        cp = self.params.to_c()
        check(lib.vc_server_create(index._h, ctypes.byref(cp), ctypes.byref(h)))
        self._h = h
        self._shapes = {}  # ticket -> (nq, k) of search requests
        self._mu = threading.Lock()

    def close(self):
        if getattr(self, "_h", None):
            check(lib.vc_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_search(self, queries, k: int) -> int:
        q = _f32(queries, self._dim)
        t = ctypes.c_uint64()
        check(lib.vc_server_submit_search(self._h, _ptr(q), S(q.shape[0]), S(int(k)), ctypes.byref(t)))
        with self._mu:
            self._shapes[t.value] = (q.shape[0], int(k))
        return t.value

    def submit_add(self, vectors) -> int:
        a = _f32(vectors, self._dim)
        t = ctypes.c_uint64()
        check(lib.vc_server_submit_add(self._h, _ptr(a), S(a.shape[0]), ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int) -> Served:
        with self._mu:
            shape = self._shapes.pop(ticket, None)
        info = RequestInfo()
        if shape is None:  # an add
            check(lib.vc_server_wait(self._h, ctypes.c_uint64(ticket), None, None, ctypes.byref(info)))
            res = None
        else:
            nq, k = shape
            scores = np.zeros(nq * k, np.float32)
            ids = np.zeros(nq * k, np.uint64)
            check(lib.vc_server_wait(self._h, ctypes.c_uint64(ticket), _ptr(scores), _ptr(ids), ctypes.byref(info)))
            ke = int(info.k_eff)
            res = SearchResults(scores[: nq * ke].reshape(nq, ke), ids[: nq * ke].reshape(nq, ke))
        return Served(res, info.t_submit, info.t_start, info.t_done, int(info.n_snapshot), int(info.k_eff),
                      int(info.batch_queries), int(info.batch), int(info.slot))

    def flush(self):
        check(lib.vc_server_flush(self._h))

    def stats(self) -> dict:
        st = ServerStats()
        check(lib.vc_server_get_stats(self._h, ctypes.byref(st)))
        return {f: int(getattr(st, f)) for f, _ in ServerStats._fields_}
