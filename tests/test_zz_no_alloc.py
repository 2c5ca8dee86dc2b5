"""soi_compress_noalloc (compress.h): the allocation-free compress(), checked against the
existing entry point compress() (called per chunk on device buffers) and against the oracle's
restatement of compress.c:205-372, word for word, on two calls in a row, the second with fewer
and shorter chunks (the static scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

from gpu_util import to_dev

pytestmark = pytest.mark.gpu


def _chunks(rng, nch, length, emi):
    x = []
    for e in emi:
        v = rng.uniform(-1, 1, length) * 2.0 ** (int(e) + 1) * 0.999
        v[rng.integers(0, length)] = 0.0
        x.append(v)
    return np.concatenate(x)


def _check(soi, orc, torch, seed, nch, length, e_max):
    L = soi.lib()
    L.soi_compress_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.soi_compress_noalloc.restype = ctypes.c_int
    rng = np.random.default_rng(seed)
    emi = np.array([e_max - int(d) for d in rng.integers(0, 40, nch)], dtype=np.int32)
    emi[0] = e_max
    x = _chunks(rng, nch, length, emi)
    stride = length * 56 // 32 + 32
    xt = to_dev(torch, x)
    wt = torch.full((nch * stride,), -1, dtype=torch.int32, device="cuda")
    nw = np.zeros(nch, dtype=np.int32)
    torch.cuda.synchronize()
    print(f"case noalloc nchunk={nch} len={length}", flush=True)
    rc = L.soi_compress_noalloc(wt.data_ptr(), stride, xt.data_ptr(), length, nch, e_max,
                                emi.ctypes.data, nw.ctypes.data, None)
    assert rc == 0, L.soi_last_error()
    w = wt.cpu().numpy().reshape(nch, stride)
    for c in range(nch):
        # the existing entry point on a device chunk and a device word buffer
        one = torch.zeros(stride, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        n = L.compress(one.data_ptr(), xt.data_ptr() + 8 * c * length, length, e_max,
                       int(emi[c]), 0)
        ref = orc.compress(x[c * length:(c + 1) * length], e_max, int(emi[c]))
        assert nw[c] == n == len(ref)
        assert np.array_equal(w[c, :n], one.cpu().numpy()[:n])
        assert np.array_equal(w[c, :n], ref)


def test_noalloc_matches_compress_two_calls(soi, orc, torch_cuda):
    _check(soi, orc, torch_cuda, 11, 9, 1000, 6)     # first call: 9 chunks, a 8-value tail each
    _check(soi, orc, torch_cuda, 12, 4, 333, -3)     # second call: fewer, shorter chunks


def test_noalloc_rejects_too_many_chunks(soi, torch_cuda):
    L = soi.lib()
    L.soi_compress_noalloc.restype = ctypes.c_int
    nw = np.zeros(4097, dtype=np.int32)
    emi = np.zeros(4097, dtype=np.int32)
    x = torch_cuda.zeros(16, dtype=torch_cuda.float64, device="cuda")
    w = torch_cuda.zeros(64, dtype=torch_cuda.int32, device="cuda")
    rc = L.soi_compress_noalloc(ctypes.c_void_p(w.data_ptr()), ctypes.c_long(64),
                                ctypes.c_void_p(x.data_ptr()), 16, 4097, 0,
                                ctypes.c_void_p(emi.ctypes.data),
                                ctypes.c_void_p(nw.ctypes.data), None)
    assert rc == 1
    assert b"4096" in L.soi_last_error()
