"""On-disk formats of the reference's recall runs (host code; no GPU needed):
NPY float32 matrices (include/vectorcache/datasets/npy.hpp:33-65,
src/datasets/npy.cpp) and the VCEX exact top-k cache
(tools/query-bench.cpp:146-340). Pinned by hand-written bytes of the
reference writers' layouts, by numpy's own NPY files, and by published FNV-1a
test vectors."""
import os
import struct

import numpy as np
import pytest

from vectorcache_amd import formats
from vectorcache_amd._capi import IndexStateError


def test_npy_writer_bytes(tmp_path):
    p = tmp_path / "sub" / "m.npy"  # parent directories are created
    x = np.arange(12, dtype=np.float32).reshape(3, 4)
    formats.write_npy_f32(p, x)
    raw = p.read_bytes()
    # version 1.0, header length 71: the 59-byte dict, 11 blanks (10 + 59 + 11
    # is a multiple of 16), a newline; then the data
    want = (b"\x93NUMPY\x01\x00G\x00{'descr': '<f4', 'fortran_order': False, 'shape': (3, 4), }" + b" " * 11 + b"\n")
    assert raw[: len(want)] == want
    assert raw[len(want):] == x.tobytes()
    assert np.array_equal(np.load(p), x)


@pytest.mark.parametrize("version", [(1, 0), (2, 0)])
def test_npy_reads_numpy_files(tmp_path, version):
    x = np.random.default_rng(1).standard_normal((37, 24)).astype(np.float32)
    p = tmp_path / "x.npy"
    with open(p, "wb") as f:
        np.lib.format.write_array(f, x, version=version)
    with formats.NpyMatrix(p) as m:
        assert (m.count, m.dim) == (37, 24)
        assert np.array_equal(m.array, x)
        assert np.array_equal(m.read_rows(30, 100), x[30:])
        assert m.read_rows(37, 5).shape == (0, 24)


def _npy(header: str, data: bytes = b"", version=(1, 0)) -> bytes:
    h = header.encode()
    if version == (1, 0):
        return b"\x93NUMPY\x01\x00" + struct.pack("<H", len(h)) + h + data
    return b"\x93NUMPY" + bytes(version) + struct.pack("<I", len(h)) + h + data


@pytest.mark.parametrize("content,msg", [
    (b"\x93NUMPY\x01", "NPY file too small"),
    (b"\x93NUMPX\x01\x00\x10\x00" + b" " * 16, "invalid NPY magic bytes"),
    (_npy("{'descr': '<f4', 'shape': (2, 2), }", version=(3, 0)), "unsupported NPY version 3.0"),
    (b"\x93NUMPY\x01\x00\xff\x00{}", "NPY header extends past end of file"),
    (_npy("{'descr': '<f8', 'fortran_order': False, 'shape': (2, 2), }", bytes(32)), "expected float32"),
    (_npy("{'descr': '<f4', 'fortran_order': False, }"), "missing 'shape' in NPY header"),
    (_npy("{'descr': '<f4', 'fortran_order': False, 'shape': (4,), }", bytes(16)), "expected 2-D NPY matrix"),
    (_npy("{'descr': '<f4', 'fortran_order': False, 'shape': (2, 2), }", bytes(15)), "NPY data truncated"),
])
def test_npy_reader_errors(tmp_path, content, msg):
    p = tmp_path / "bad.npy"
    p.write_bytes(content)
    with pytest.raises(IndexStateError, match=msg):
        formats.NpyMatrix(p)


def test_npy_reader_accepts_double_quoted_descr_and_missing_file(tmp_path):
    p = tmp_path / "dq.npy"
    x = np.arange(6, dtype=np.float32)
    p.write_bytes(_npy('{"descr": "<f4", "fortran_order": False, "shape": (2, 3), }\n', x.tobytes()))
    with formats.NpyMatrix(p) as m:
        assert np.array_equal(m.array, x.reshape(2, 3))
    with pytest.raises(IndexStateError, match="failed to open"):
        formats.NpyMatrix(tmp_path / "missing.npy")


KEY = dict(source_label="npy-openai-1f", split="train", query_split="test", index_n=1000, seed=0)


def test_vcex_bytes_and_round_trip(tmp_path):
    ids = np.array([[5, 1], [7, 3]], np.uint64)
    p = tmp_path / "c" / "x.bin"
    formats.vcex_save(p, ids, **KEY)
    raw = p.read_bytes()
    want = (b"VCEX" + struct.pack("<I", 1) + struct.pack("<4Q", 2, 2, 1000, 0) +
            struct.pack("<I", 13) + b"npy-openai-1f" + struct.pack("<I", 5) + b"train" +
            struct.pack("<I", 4) + b"test" + struct.pack("<4Q", 5, 1, 7, 3))
    assert raw == want
    assert not (tmp_path / "c" / "x.bin.tmp").exists()
    got = formats.vcex_load(p, query_n=2, k=2, **KEY)
    assert np.array_equal(got, ids)


@pytest.mark.parametrize("change", [dict(index_n=999), dict(seed=1), dict(split="test"),
                                    dict(query_split="train"), dict(source_label="npy-other-1f")])
def test_vcex_key_mismatch_is_a_miss(tmp_path, change):
    p = tmp_path / "x.bin"
    formats.vcex_save(p, np.zeros((3, 4), np.uint64), **KEY)
    assert formats.vcex_load(p, query_n=3, k=4, **{**KEY, **change}) is None
    assert formats.vcex_load(p, query_n=3, k=5, **KEY) is None
    assert formats.vcex_load(p, query_n=2, k=4, **KEY) is None
    assert formats.vcex_load(tmp_path / "missing.bin", query_n=3, k=4, **KEY) is None


def test_vcex_bad_magic_or_version_is_a_miss(tmp_path):
    p = tmp_path / "x.bin"
    formats.vcex_save(p, np.zeros((1, 1), np.uint64), **KEY)
    raw = bytearray(p.read_bytes())
    q = tmp_path / "v2.bin"
    q.write_bytes(bytes(raw[:4]) + struct.pack("<I", 2) + bytes(raw[8:]))
    assert formats.vcex_load(q, query_n=1, k=1, **KEY) is None
    q.write_bytes(b"VCEY" + bytes(raw[4:]))
    assert formats.vcex_load(q, query_n=1, k=1, **KEY) is None
    q.write_bytes(bytes(raw[:-1]))  # truncated ids
    assert formats.vcex_load(q, query_n=1, k=1, **KEY) is None


def test_vcex_source_label_and_cache_path():
    # FNV-1a 64 published vectors: "foobar" -> 85944171f73967e8, "a" -> af63dc4c8601ec8c
    assert formats.vcex_source_label("foobar") == "npy-foobar-85944171f73967e8"
    assert formats.vcex_source_label("a") == "npy-a-af63dc4c8601ec8c"
    assert formats.vcex_source_label("", "openai-1536") == "openai-1536"
    path = formats.vcex_cache_path(source_label="npy-a b", split="train", query_split="te/st", index_n=10,
                                   query_n=2, k=3, seed=7)
    assert path == os.path.join(".cache", "exact_topk", "npy-a_b_train_n10_qs-te_st_ql2_k3_seed7.bin")
