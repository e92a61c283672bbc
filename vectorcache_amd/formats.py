"""On-disk formats of the reference's recall runs (host code, no device):
NPY float32 matrices through mmap (NpyReader, include/vectorcache/datasets/
npy.hpp:33-65; write_npy_f32_matrix, src/datasets/npy.cpp:336-366) and the
VCEX exact top-k cache of query-bench (tools/query-bench.cpp:146-340)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._capi import P, S, VcexKey, check, lib


def _b(s) -> bytes:
    return os.fsencode(s) if not isinstance(s, bytes) else s


class NpyMatrix:
    """A read-only mapping of a float32 (count, dim) NPY file. `array` is a
    zero-copy numpy view of the rows (valid while the object lives)."""

    def __init__(self, path):
        h, n, d = P(), S(), S()
        check(lib.vc_npy_open(_b(path), ctypes.byref(h), ctypes.byref(n), ctypes.byref(d)))
        self._h = h
        self.count, self.dim = n.value, d.value
        if self.count * self.dim > 0:
            ptr = lib.vc_npy_data(self._h)
            self.array = np.ctypeslib.as_array(ptr, shape=(self.count, self.dim))
            self.array.flags.writeable = False
        else:
            self.array = np.zeros((self.count, self.dim), np.float32)

    def read_rows(self, first: int, n: int) -> np.ndarray:
        """Copy of rows [first, first + n) clipped to the row count."""
        out = np.zeros((max(0, min(n, self.count - first)), self.dim), np.float32)
        got = S()
        check(lib.vc_npy_read_rows(self._h, S(first), S(n), out.ctypes.data_as(P), ctypes.byref(got)))
        return out[: got.value]

    def close(self) -> None:
        if self._h:
            self.array = None
            lib.vc_npy_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_npy_f32(path, matrix) -> None:
    """write_npy_f32_matrix: the reference writer's bytes (version 1.0, header
    padded to 16 bytes before its newline)."""
    m = np.ascontiguousarray(matrix, np.float32)
    if m.ndim != 2:
        raise ValueError("expected a 2-D matrix")
    check(lib.vc_npy_write_f32(_b(path), S(m.shape[0]), S(m.shape[1]), m.ctypes.data_as(P)))


def _key(source_label, split, query_split, index_n, query_n, k, seed=0) -> VcexKey:
    return VcexKey(_b(source_label), _b(split), _b(query_split), int(index_n), int(query_n), int(k), int(seed))


def vcex_save(path, ids, *, source_label, split, query_split, index_n, seed=0) -> None:
    """Write query_n x k exact top-k ids as a VCEX cache file."""
    ids = np.ascontiguousarray(ids, np.uint64)
    if ids.ndim != 2:
        raise ValueError("ids must be query_n x k")
    key = _key(source_label, split, query_split, index_n, ids.shape[0], ids.shape[1], seed)
    check(lib.vc_vcex_save(_b(path), ctypes.byref(key), ids.ctypes.data_as(P)))


def vcex_load(path, *, source_label, split, query_split, index_n, query_n, k, seed=0):
    """The cached ids (query_n x k u64), or None when the file is missing,
    not VCEX version 1 or made for another key."""
    key = _key(source_label, split, query_split, index_n, query_n, k, seed)
    ids = np.zeros((int(query_n), int(k)), np.uint64)
    loaded = ctypes.c_int()
    check(lib.vc_vcex_load(_b(path), ctypes.byref(key), ids.ctypes.data_as(P), ctypes.byref(loaded)))
    return ids if loaded.value else None


def _string(call) -> str:
    n = S()
    check(call(None, S(0), ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    check(call(buf, S(n.value + 1), ctypes.byref(n)))
    return os.fsdecode(buf.value)


def vcex_source_label(npy_path, dataset: str = "") -> str:
    """cache_source_label: "npy-<stem>-<FNV-1a 64 hex>" or the dataset name."""
    return _string(lambda out, cap, n: lib.vc_vcex_source_label(_b(npy_path), _b(dataset), out, cap, n))


def vcex_cache_path(*, source_label, split, query_split, index_n, query_n, k, seed=0) -> str:
    """exact_topk_cache_path: .cache/exact_topk/<key>.bin (relative)."""
    key = _key(source_label, split, query_split, index_n, query_n, k, seed)
    return _string(lambda out, cap, n: lib.vc_vcex_cache_path(ctypes.byref(key), out, cap, n))
