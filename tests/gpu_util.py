"""Helpers for GPU tests: device memory via torch (plumbing), raw copies via the HIP runtime."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    return _hip


def d2h(ptr, count, dtype):
    out = np.empty(count, dtype=dtype)
    rc = hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)  # hipMemcpyDeviceToHost
    assert rc == 0, rc
    return out


def to_dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
