"""The search entry point that allocates nothing (vc_index_search_device_noalloc)
against vc_index_search_device, called through the C ABI: bit-identical
results on two calls in a row (the second with fewer queries and a smaller k),
and a request larger than the fixed scratch rejected with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, lib, ix, q, k, s, i, stream):
    P, S = ctypes.c_void_p, ctypes.c_size_t
    kout = S()
    st = getattr(lib, fn)(ix._h, P(q.data_ptr()), S(q.shape[0]), S(k), P(s.data_ptr()), P(i.data_ptr()),
                          ctypes.byref(kout), P(stream))
    return st, kout.value


def test_noalloc_search_matches_search_device():
    import torch

    from vectorcache_amd import synthetic
    from vectorcache_amd._capi import VC_ERR_INVALID_ARGUMENT, VC_OK, lib
    from vectorcache_amd.index import TurboQuantIndex

    dev = torch.device("cuda:0")
    x = synthetic.anisotropic(30000, 1536, 91)
    q = synthetic.anisotropic(64, 1536, 92)
    ix = TurboQuantIndex(1536, 4)
    ix.calibrate(synthetic.anisotropic(2000, 1536, 93))
    ix.add(x)
    ix.prepare()
    qd = torch.from_numpy(q).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    assert lib.vc_noalloc_scratch_bytes() == 16 << 20
    for nq, k in ((64, 100), (5, 10)):
        qq = qd[:nq].contiguous()
        s0 = torch.empty((nq, k), dtype=torch.float32, device=dev)
        i0 = torch.empty((nq, k), dtype=torch.int64, device=dev)
        s1 = torch.full((nq, k), 7.0, dtype=torch.float32, device=dev)
        i1 = torch.full((nq, k), 7, dtype=torch.int64, device=dev)
        st0, k0 = _call("vc_index_search_device", lib, ix, qq, k, s0, i0, stream)
        st1, k1 = _call("vc_index_search_device_noalloc", lib, ix, qq, k, s1, i1, stream)
        assert st0 == VC_OK and st1 == VC_OK and k0 == k1 == k
        torch.cuda.synchronize()
        assert torch.equal(i0, i1), (nq, k)
        assert np.array_equal(s0.cpu().numpy().view(np.uint32), s1.cpu().numpy().view(np.uint32)), (nq, k)
    # 4096 queries need about 100 MB of LUTs alone: rejected, outputs untouched
    big = torch.from_numpy(synthetic.anisotropic(4096, 1536, 94)).to(dev)
    sb = torch.full((4096, 10), 3.0, dtype=torch.float32, device=dev)
    ib = torch.full((4096, 10), 3, dtype=torch.int64, device=dev)
    st, _ = _call("vc_index_search_device_noalloc", lib, ix, big, 10, sb, ib, stream)
    assert st == VC_ERR_INVALID_ARGUMENT
    assert b"fixed scratch" in lib.vc_last_error()
    torch.cuda.synchronize()
    assert bool((sb == 3.0).all()) and bool((ib == 3).all())
