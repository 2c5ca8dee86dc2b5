"""The reference's test-harness entry points exported by the library (soi.h:282-288, 306,
314-315): input kinds 0, 1, 3 and their closed-form transforms (input.c:218-290), SNR and the
normalised infinity norm (input.c:292-425), get_cpu_freq and load_imbalance_times.  Host code
only, no GPU.  The closed forms are checked against numpy's FFT of the generated inputs (an
independent FP64 DFT), so a wrong sign, index or scale in either fails."""
import ctypes

import numpy as np
import pytest

import soi_fft_amd as soi

dp = ctypes.POINTER(ctypes.c_double)


class C128(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


@pytest.fixture(scope="module")
def L():
    lib = soi.lib()
    lib.populate_input.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                   ctypes.c_size_t, ctypes.c_int]
    lib.populate_input.restype = None
    lib.reference_output.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_size_t]
    lib.reference_output.restype = C128
    lib.compute_snr.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    lib.compute_snr.restype = ctypes.c_double
    lib.compute_normalized_inf_norm.argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
    lib.compute_normalized_inf_norm.restype = ctypes.c_double
    lib.get_cpu_freq.restype = ctypes.c_double
    return lib


def populate(L, n, kind, local=None, offset=0):
    local = n if local is None else local
    buf = np.zeros(local, dtype=np.complex128)
    L.populate_input(buf.ctypes.data, local, offset, n, kind)
    return buf


def reference(L, n, kind):
    out = np.empty(n, dtype=np.complex128)
    for i in range(n):
        r = L.reference_output(i, n, kind, 0)
        out[i] = r.re + 1j * r.im
    return out


@pytest.mark.parametrize("n", [1024, 3 * 512])
@pytest.mark.parametrize("kind", [0, 1])
def test_closed_forms_match_numpy_fft(L, n, kind):
    x = populate(L, n, kind)
    ref = reference(L, n, kind)
    got = np.fft.fft(x)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12


@pytest.mark.parametrize("n", [1024, 3 * 512])
def test_kind3_reference_is_transform_over_n(L, n):
    """Kind 3 (Dirichlet kernel) keeps the reference's expected value, the rectangle of height
    1/W, which is the exact DFT divided by N."""
    x = populate(L, n, 3)
    ref = reference(L, n, 3)
    got = np.fft.fft(x) / n
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12


def test_inputs_are_the_documented_functions(L):
    n = 1024
    W = n // 16 + 1
    idx = np.arange(n)
    rect = np.where((2 * idx < W) | (2 * (n - idx) < W), 1.0 / W, 0.0)
    assert np.array_equal(populate(L, n, 0), rect.astype(np.complex128))
    np.testing.assert_allclose(populate(L, n, 1), np.exp(1j * idx), rtol=0, atol=1e-13)
    x3 = populate(L, n, 3)
    assert x3[0] == 1
    k = idx[1:]
    np.testing.assert_allclose(x3[1:].real, np.sin(np.pi * W * k / n) / (W * np.sin(np.pi * k / n)),
                               rtol=0, atol=1e-12)


@pytest.mark.parametrize("kind", [0, 1, 3])
def test_local_window_equals_slice(L, kind):
    n, off, loc = 4096, 1000, 777
    full = populate(L, n, kind)
    part = populate(L, n, kind, local=loc, offset=off)
    assert np.array_equal(part, full[off:off + loc])


@pytest.mark.parametrize("kind", [0, 1])
def test_snr_and_inf_norm(L, kind):
    n = 2048
    y = np.fft.fft(populate(L, n, kind))
    snr = L.compute_snr(y.ctypes.data, n, 0, n, kind, None)
    inf = L.compute_normalized_inf_norm(y.ctypes.data, n, 0, n, kind)
    assert snr > 250 and inf < 1e-12
    # a planted error of relative size 1e-6 at one bin lowers both as the definitions say
    z = y.copy()
    ref = reference(L, n, kind)
    j = int(np.argmax(np.abs(ref)))
    z[j] += 1e-6 * np.abs(ref).max()
    inf2 = L.compute_normalized_inf_norm(z.ctypes.data, n, 0, n, kind)
    snr2 = L.compute_snr(z.ctypes.data, n, 0, n, kind, None)
    assert abs(inf2 - 1e-6) < 1e-9
    expect = 10 * np.log10(np.sum(np.abs(ref) ** 2) / (1e-6 * np.abs(ref).max()) ** 2)
    assert abs(snr2 - expect) < 0.01


@pytest.mark.parametrize("kind", [2, 4, 5, 9])
def test_unsupported_kinds_report_errors(L, kind):
    buf = np.full(16, 7 + 7j, dtype=np.complex128)
    L.populate_input(buf.ctypes.data, 16, 0, 16, kind)
    assert soi.lib().soi_last_status() == 1
    assert np.all(buf == 7 + 7j)
    assert np.isnan(L.compute_snr(buf.ctypes.data, 16, 0, 16, kind, None))
    L.populate_input(buf.ctypes.data, 16, 0, 16, 1)
    assert soi.lib().soi_last_status() == 0


def test_cpu_freq_and_imbalance_counters(L):
    f = L.get_cpu_freq()
    assert 1e8 < f < 1e11
    arr = (ctypes.c_ulonglong * 512).in_dll(L, "load_imbalance_times")
    assert all(v == 0 for v in arr)
