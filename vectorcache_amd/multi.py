"""Many collections searched together (vc_multi_* of the C ABI), and the
scan-plan autotuner (vc_tune_*).

A MultiSearch holds prepared TurboQuantIndex objects of any dim / bits / size,
with a query batch size and k per collection fixed at construction. Its
chain (query preparation per collection, one persistent scan launch over
work items of every collection, the exact merge per collection) runs eagerly
or as a captured HIP graph; every collection's results are bit-identical to
its own search_device of the same queries. No counterpart in the reference,
which searches one index per call (src/index.cpp:267-289).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from ._capi import MultiPlanInfo, P, S, ScanChoice, check, lib
from .index import TurboQuantIndex, _cuda_rows, _launch_stream


@dataclass
class MultiPlan:
    tile: int
    vec_per_thread: int
    item_scale: float
    workgroups_per_cu: int
    items: int
    grid: int
    lds_bytes: int
    measured_ms: float


def _plan(info: MultiPlanInfo) -> MultiPlan:
    return MultiPlan(int(info.tile), int(info.vec_per_thread), float(info.item_scale), int(info.workgroups_per_cu),
                     int(info.items), int(info.grid), int(info.lds_bytes), float(info.measured_ms))


class MultiSearch:
    def __init__(self, indexes: list[TurboQuantIndex], nq: list[int], k: list[int]):
        import torch

        if not (len(indexes) == len(nq) == len(k)) or not indexes:
            raise ValueError("indexes, nq and k must be non-empty lists of one length")
        self.indexes = list(indexes)
        self.nq = [int(v) for v in nq]
        self.k = [int(v) for v in k]
        arr_i = (P * len(indexes))(*(ix._h for ix in indexes))
        arr_q = (S * len(indexes))(*self.nq)
        arr_k = (S * len(indexes))(*self.k)
        h = P()
        check(lib.vc_multi_create(ctypes.cast(arr_i, P), S(len(indexes)), ctypes.cast(arr_q, P),
                                  ctypes.cast(arr_k, P), ctypes.byref(h)))
        self._h = h
        self._rows = []
        for j, ix in enumerate(indexes):
            q, sc, ids, ke = P(), P(), P(), S()
            check(lib.vc_multi_rows(h, S(j), ctypes.byref(q), ctypes.byref(sc), ctypes.byref(ids), ctypes.byref(ke)))
            self._rows.append((q.value, sc.value, ids.value, int(ke.value), ix.dim()))
        self._dev = torch.device("cuda", torch.cuda.current_device())

    def close(self):
        if getattr(self, "_h", None):
            check(lib.vc_multi_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_queries(self, j: int, queries, stream=None):
        """Copy nq[j] x dim queries (a CUDA tensor) into collection j's rows,
        ordered on `stream` (default: torch's current stream)."""
        q = _cuda_rows(queries, self._rows[j][4], "queries")
        if q.numel() != self.nq[j] * self._rows[j][4]:
            raise ValueError(f"collection {j} takes {self.nq[j]} queries")
        check(lib.vc_multi_set_queries(self._h, S(j), P(q.data_ptr()), _launch_stream(stream, q)))

    def results(self, j: int, stream=None):
        """Collection j's (scores f32, ids int64 holding the u64 ids), copied
        out on `stream` (default: torch's current stream)."""
        import torch

        ke = self._rows[j][3]
        nq = self.nq[j]
        s = torch.empty((nq, ke), dtype=torch.float32, device=self._dev)
        i = torch.empty((nq, ke), dtype=torch.int64, device=self._dev)
        check(lib.vc_multi_get_results(self._h, S(j), P(s.data_ptr()), P(i.data_ptr()), _launch_stream(stream, s, i)))
        return s, i

    def search(self, stream=None):
        check(lib.vc_multi_search(self._h, _launch_stream(stream)))

    def capture(self, stream):
        check(lib.vc_multi_capture(self._h, P(int(stream))))

    def replay(self, stream=None):
        check(lib.vc_multi_replay(self._h, _launch_stream(stream)))

    def set_plan(self, tile: int, vec_per_thread: int, item_scale: float = 1.0, workgroups_per_cu: int = 2):
        check(lib.vc_multi_set_plan(self._h, S(tile), S(vec_per_thread), ctypes.c_float(item_scale),
                                    S(workgroups_per_cu)))

    def plan(self) -> MultiPlan:
        info = MultiPlanInfo()
        check(lib.vc_multi_get_plan(self._h, ctypes.byref(info)))
        return _plan(info)

    def collection_plan(self, j: int) -> tuple[int, int]:
        """(query tile, vector ranges) of collection j under the plan."""
        t, l = S(), S()
        check(lib.vc_multi_collection_plan(self._h, S(j), ctypes.byref(t), ctypes.byref(l)))
        return int(t.value), int(l.value)

    def tune(self, reps: int = 3) -> MultiPlan:
        info = MultiPlanInfo()
        check(lib.vc_multi_tune(self._h, S(reps), ctypes.byref(info)))
        return _plan(info)


def tune_scan(dim: int, bits: int, n: int, nq: int, k: int, reps: int = 3) -> dict:
    """Measure the single-index scan plans of a shape and record the fastest."""
    c = ScanChoice()
    check(lib.vc_tune_scan(S(dim), S(bits), S(n), S(nq), S(k), S(reps), ctypes.byref(c)))
    return {"tile": int(c.tile), "vec_per_thread": int(c.vec_per_thread), "range_scale": float(c.range_scale),
            "ms": float(c.ms), "rule_ms": float(c.rule_ms), "rule_tile": int(c.rule_tile)}


def save_tuning(path) -> None:
    check(lib.vc_tune_save(str(path).encode()))


def load_tuning(path) -> int:
    n = S()
    check(lib.vc_tune_load(str(path).encode(), ctypes.byref(n)))
    return int(n.value)


def clear_tuning() -> None:
    check(lib.vc_tune_clear())


def tuning_entries() -> int:
    return int(lib.vc_tune_count())
