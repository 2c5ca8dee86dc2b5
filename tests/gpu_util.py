"""Helpers for GPU tests: device memory via torch (plumbing), raw copies via the HIP runtime,
and the relative-L2 helper, which records every achieved error under the running test's id so
that the session ends with a table of them (conftest.pytest_terminal_summary)."""
import ctypes
import os

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


_RECORD = {}


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    r = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    tid = os.environ.get("PYTEST_CURRENT_TEST", "?").rsplit(" (", 1)[0]
    _RECORD.setdefault(tid, []).append(r)
    return r


def rel_t(torch, a, b):
    """rel() for device tensors, computed on the device (full-size outputs) and recorded too."""
    r = float((torch.linalg.norm(a - b) / torch.linalg.norm(b)).item())
    tid = os.environ.get("PYTEST_CURRENT_TEST", "?").rsplit(" (", 1)[0]
    _RECORD.setdefault(tid, []).append(r)
    return r


def recorded():
    """{test id: [relative L2 errors of its comparisons, in order]}"""
    return _RECORD


def hip_malloc(nbytes):
    """A plain hipMalloc allocation (its own address range, unlike torch's caching allocator)."""
    h = hip()
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMalloc.restype = ctypes.c_int
    p = ctypes.c_void_p()
    rc = h.hipMalloc(ctypes.byref(p), nbytes)
    assert rc == 0, rc
    return p.value


def hip_free(ptr):
    h = hip()
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipFree.restype = ctypes.c_int
    assert h.hipFree(ptr) == 0
