"""ctypes binding of the CPU oracle (oracle/soi_oracle.c) and of the CPU performance port
(oracle/soi_cpu_port.c, the CPU baseline bench.py times).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module as the checker and the baseline.  The product package (soi_fft_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_long),
        ("P", ctypes.c_int),
        ("k", ctypes.c_int),
        ("n_mu", ctypes.c_int),
        ("d_mu", ctypes.c_int),
        ("B", ctypes.c_int),
        ("tau", ctypes.c_double),
        ("sigma", ctypes.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        pp = ctypes.POINTER(OrcParams)
        _lib.orc_w_table.argtypes = [pp, dp]
        _lib.orc_w_inv_table.argtypes = [pp, dp]
        _lib.orc_filter_subsample.argtypes = [pp, dp, dp, dp]
        _lib.orc_filter_brute.argtypes = [pp, dp, dp]
        _lib.orc_fft.argtypes = [ctypes.c_long, dp]
        _lib.orc_fft_batch.argtypes = [ctypes.c_long, ctypes.c_long, dp]
        _lib.orc_dft_brute.argtypes = [ctypes.c_long, dp, dp]
        _lib.orc_demodulate.argtypes = [pp, dp, dp, dp]
        _lib.orc_compute_soi.argtypes = [pp, dp, ctypes.c_int, ctypes.c_double, ip]
        _lib.orc_compute_soi.restype = ctypes.c_int
        _lib.orc_max_exponent.argtypes = [dp, ctypes.c_int]
        _lib.orc_max_exponent.restype = ctypes.c_int
        _lib.orc_compress_mode.argtypes = [ip, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int]
        _lib.orc_compress_mode.restype = ctypes.c_int
        _lib.orc_decompress.argtypes = [dp, ip, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib.orc_vlc_plan.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_long,
                                      ctypes.c_double, ip, ip, ip]
        _lib.orc_uniform_input.argtypes = [ctypes.c_long, ctypes.c_uint64, dp]
        _lib.orc_num_threads.restype = ctypes.c_int
        _lib.orc_set_window_mode.argtypes = [ctypes.c_int]
        _lib.port_create.argtypes = [pp]
        _lib.port_create.restype = ctypes.c_void_p
        _lib.port_compute.argtypes = [ctypes.c_void_p, dp]
        _lib.port_destroy.argtypes = [ctypes.c_void_p]
        _lib.port_threads.restype = ctypes.c_int
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def params(N, P, k=1, n_mu=5, d_mu=4, B=72, tau=928 / 1024.0, sigma=373.7314):
    """Defaults are set_default_soi_descriptor's (pfft.c:87-91)."""
    return OrcParams(N, P, k, n_mu, d_mu, B, tau, sigma)


def dims(p):
    S = p.k * p.P
    M = p.N // S
    Mh = p.n_mu * M // p.d_mu
    return S, M, Mh


def as_f64(c):
    return np.ascontiguousarray(np.asarray(c, dtype=np.complex128)).view(np.float64)


def set_window_mode(mode):
    """0: windows exactly as the reference forms t (default); 1: exact rational t."""
    lib().orc_set_window_mode(mode)


def w_table(p):
    S, M, Mh = dims(p)
    w = np.zeros(2 * p.n_mu * p.B * S)
    lib().orc_w_table(ctypes.byref(p), _dp(w))
    return w.view(np.complex128).reshape(p.n_mu, p.B, S)


def w_inv_table(p):
    S, M, Mh = dims(p)
    w = np.zeros(2 * M)
    lib().orc_w_inv_table(ctypes.byref(p), _dp(w))
    return w.view(np.complex128)


def filter_subsample(p, x):
    S, M, Mh = dims(p)
    w = as_f64(w_table(p).reshape(-1))
    xi = as_f64(x)
    out = np.zeros(2 * S * Mh)
    lib().orc_filter_subsample(ctypes.byref(p), _dp(w), _dp(xi), _dp(out))
    return out.view(np.complex128).reshape(S, Mh)


def filter_brute(p, x):
    S, M, Mh = dims(p)
    xi = as_f64(x)
    out = np.zeros(2 * S * Mh)
    lib().orc_filter_brute(ctypes.byref(p), _dp(xi), _dp(out))
    return out.view(np.complex128).reshape(S, Mh)


def fft(x):
    a = as_f64(x).copy()
    lib().orc_fft(len(a) // 2, _dp(a))
    return a.view(np.complex128)


def fft_batch(x2d):
    a = as_f64(x2d).copy()
    n, L = x2d.shape
    lib().orc_fft_batch(L, n, _dp(a))
    return a.view(np.complex128).reshape(n, L)


def dft_brute(x):
    a = as_f64(x)
    out = np.zeros_like(a)
    lib().orc_dft_brute(len(a) // 2, _dp(a), _dp(out))
    return out.view(np.complex128)


def compute_soi(p, x, use_vlc=False, comm_to_comp=1.0):
    a = as_f64(x).copy()
    sb = np.zeros(p.P + 1, dtype=np.int32)
    rc = lib().orc_compute_soi(ctypes.byref(p), _dp(a), int(use_vlc), comm_to_comp, _ip(sb))
    assert rc == 0
    return a.view(np.complex128), sb


def max_exponent(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_max_exponent(_dp(a), len(a))


def compress(x, e_max, e_max_i, floor_mode=False):
    a = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(2 * len(a) + 64, dtype=np.int32)
    n = lib().orc_compress_mode(_ip(out), _dp(a), len(a), e_max, e_max_i, int(floor_mode))
    return out[:n].copy()


def decompress(words, length, e_max, e_max_i):
    w = np.zeros(2 * length + 64, dtype=np.int32)
    w[: len(words)] = words
    out = np.zeros(length)
    lib().orc_decompress(_dp(out), _ip(w), length, e_max, e_max_i)
    return out


def vlc_plan(P, k, l, M_hat, comm_to_comp, gmax):
    g = np.ascontiguousarray(gmax, dtype=np.int32)
    sc = np.zeros(P * k, dtype=np.int32)
    sb = np.zeros(P + 1, dtype=np.int32)
    lib().orc_vlc_plan(P, k, l, M_hat, comm_to_comp, _ip(g), _ip(sc), _ip(sb))
    return sc, sb


def uniform_input(N, seed):
    a = np.zeros(2 * N)
    lib().orc_uniform_input(N, seed, _dp(a))
    return a.view(np.complex128)


class CpuPort:
    """The CPU performance port of compute_soi (soi_cpu_port.c): plan once, transform in place.
    A baseline to time beside the GPU, not the checker."""

    def __init__(self, p):
        self._p = p
        self.h = lib().port_create(ctypes.byref(p))
        if not self.h:
            raise ValueError("port_create: unsupported shape")

    def compute(self, x):
        """x: complex128 array of N values, transformed in place."""
        assert x.dtype == np.complex128 and x.flags["C_CONTIGUOUS"] and x.size == self._p.N
        lib().port_compute(self.h, x.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return x

    def close(self):
        if self.h:
            lib().port_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def port_threads():
    return lib().port_threads()
