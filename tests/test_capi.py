"""The C ABI library without a GPU: it builds for gfx950, loads, exports every
symbol include/vectorcache_hip.h declares, its host precompute (rotation
tables, codebook) is bit-identical to the reference's golden vectors, and
compute entry points fail loudly when no device is present."""
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, load_golden
from util import bits_equal

HEADER = ROOT / "include" / "vectorcache_hip.h"


@pytest.fixture(scope="module")
def capi():
    so = ROOT / "vectorcache_amd" / "libvectorcache_hip.so"
    if not so.exists():
        subprocess.run([sys.executable, "-m", "vectorcache_amd.build"], cwd=ROOT, check=True)
    from vectorcache_amd import _capi

    return _capi


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for must in ("vc_index_create", "vc_index_calibrate", "vc_index_add", "vc_index_prepare", "vc_index_search",
                 "vc_encode", "vc_search_perm0", "vc_merge_results", "vc_last_error"):
        assert must in syms


def test_library_exports_every_declared_symbol(capi):
    missing = [s for s in declared_symbols() if not hasattr(capi.lib, s)]
    assert not missing, missing
    assert set(declared_symbols()) <= set(capi.EXPORTED)


def test_library_is_built_for_gfx950(capi):
    assert "gfx950" in capi.version()
    blob = (ROOT / "vectorcache_amd" / "libvectorcache_hip.so").read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob  # the embedded code object's target


def test_host_precompute_matches_golden(capi):
    import vectorcache_amd.index as vc

    g = load_golden("rotation.npz")
    for key in g:
        if key.startswith("perm_"):
            d = int(key.split("_")[1])
            perm, signs, isb, blk = vc.rotation_tables(d)
            assert np.array_equal(perm, g[key]) and bits_equal(signs, g[f"signs_{d}"]), d
            assert bits_equal(np.float32([isb]), g[f"isb_{d}"]) and blk == int(g[f"block_{d}"][0])
    cb = load_golden("codebook.npz")
    for key in cb:
        if key.startswith("cen_"):
            _, bits, d = key.split("_")
            b, c = vc.codebook(int(bits), int(d))
            assert bits_equal(c, cb[key]) and bits_equal(b, cb[f"bnd_{bits}_{d}"]), key


def test_invalid_arguments_map_to_value_error(capi):
    import vectorcache_amd.index as vc

    with pytest.raises(ValueError):
        vc.codebook(5, 128)
    with pytest.raises(ValueError):
        vc.rotation_tables(100)
    with pytest.raises(ValueError):
        vc.set_scan_tile(3)


def _gpu_present():
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_present(), reason="checks the no-device error path")
def test_compute_fails_loudly_without_a_device(capi):
    import vectorcache_amd.index as vc

    with pytest.raises(capi.VectorCacheError):
        vc.TurboQuantIndex(128, 4)
    with pytest.raises(capi.VectorCacheError):
        vc.encode(np.ones((4, 128), np.float32), 4)


def test_score_into_argument_errors_without_device(capi):
    """score_prepared_into's argument checks (the reference's exceptions) come
    before any device work: empty query_indices -> ValueError
    (std::invalid_argument), an index >= nq -> IndexError (std::out_of_range),
    n == 0 or k == 0 -> no change."""
    import vectorcache_amd.index as vc

    q = np.zeros((3, 64), np.float32)
    heaps = vc.empty_heaps(3, 5)
    with pytest.raises(ValueError):
        vc.score_perm0_into(q, 5, 4, np.zeros(0, np.uint8), np.zeros(0, np.float32), [], heaps)
    with pytest.raises(IndexError):
        vc.score_perm0_into(q, 5, 4, np.zeros(0, np.uint8), np.zeros(0, np.float32), [0, 3], heaps)
    before = [h.copy() for h in heaps]
    vc.score_perm0_into(q, 5, 4, np.zeros(0, np.uint8), np.zeros(0, np.float32), [2, 0], heaps)  # n == 0
    assert all(np.array_equal(a, b) for a, b in zip(before, heaps))
    heaps[4][1] = np.uint64(2**64 - 1)  # the buffered collector marker
    with pytest.raises(ValueError):
        vc.score_perm0_into(q, 5, 4, np.zeros(32 * 192, np.uint8), np.ones(1, np.float32), [1], heaps)
    heaps[4][1] = 0
    with pytest.raises(ValueError):  # a repeated index
        vc.score_perm0_into(q, 5, 4, np.zeros(32 * 192, np.uint8), np.ones(1, np.float32), [1, 1], heaps)
