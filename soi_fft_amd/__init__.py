"""soi_fft_amd -- SOI FFT for AMD Instinct MI355X behind the soi.h C ABI.

The product is libsoi_hip.so (C host layer + HIP kernels, built by soi_fft_amd.build).  This
module is a thin ctypes mirror of the C API (reference soi.h:236-288) so Python callers and the
tests drive exactly the entry points a C caller would:

    d = SoiDesc(); set_default_soi_descriptor(d)
    d.N, d.P = 1 << 27, 32
    init_soi_descriptor(d, k=1)
    compute_soi(d, ptr)            # ptr: device (or host) address of N complex128
    free_soi_descriptor(d)

There is no fallback: if the library is missing or a call fails, an exception is raised.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsoi_hip.so")
_lib = None

SOI_PREC_FP64 = 0
SOI_PREC_FP32 = 1
SOI_WINDOW_AUTO = 0
SOI_WINDOW_REFERENCE = 1
SOI_WINDOW_EXACT = 2


# status codes (soi.h)
SOI_OK, SOI_ERR_ARGS, SOI_ERR_TOO_SMALL, SOI_ERR_UNSUPPORTED_MU = 0, 1, 2, 3
SOI_ERR_ALLOC, SOI_ERR_HIP, SOI_ERR_STATE = 4, 5, 6


class SoiError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"soi error {status}: {msg}")
        self.status = status


class SoiDesc(ctypes.Structure):
    """Layout of soi_desc_t in include/soi.h (reference soi.h:236-270)."""
    _fields_ = [
        ("comm", ctypes.c_int64),
        ("P", ctypes.c_int),
        ("rank", ctypes.c_int),
        ("k", ctypes.c_size_t),
        ("N", ctypes.c_size_t),
        ("W_inv", ctypes.c_void_p),
        ("w", ctypes.c_void_p),
        ("w_dup", ctypes.c_void_p),
        ("gamma_tilde", ctypes.c_void_p),
        ("alpha_tilde", ctypes.c_void_p),
        ("beta_tilde", ctypes.c_void_p),
        ("alpha_ghost", ctypes.c_void_p),
        ("delta", ctypes.c_void_p),
        ("epsilon", ctypes.c_void_p),
        ("n_mu", ctypes.c_int),
        ("d_mu", ctypes.c_int),
        ("B", ctypes.c_size_t),
        ("tau", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("plan", ctypes.c_void_p),
        ("sendRequests_", ctypes.c_void_p),
        ("recvRequests_", ctypes.c_void_p),
        ("use_vlc", ctypes.c_int),
        ("segmentBoundaries", ctypes.POINTER(ctypes.c_int)),
        ("comm_to_comp_cost_ratio", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("window_mode", ctypes.c_int),
    ]

    # derived sizes (pfft.c:280-283)
    @property
    def S(self):
        return self.k * self.P

    @property
    def M(self):
        return self.N // self.S

    @property
    def M_hat(self):
        return self.n_mu * self.M // self.d_mu


def lib():
    """Load libsoi_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with python -m soi_fft_amd.build")
        # A process must hold one HIP runtime.  PyTorch ships its own libamdhip64 (same soname,
        # different file); when torch is present it is loaded first so that libsoi_hip.so binds
        # to that runtime instead of loading a second one from /opt/rocm.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(SoiDesc)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.set_default_soi_descriptor.argtypes = [P]
        L.init_soi_descriptor.argtypes = [P, ctypes.c_int, ctypes.c_size_t]
        L.compute_soi.argtypes = [P, vp]
        L.compute_soi_async.argtypes = [P, vp, vp]
        L.free_soi_descriptor.argtypes = [P]
        L.parallel_filter_subsampling.argtypes = [P, vp]
        L.soi_last_error.restype = ctypes.c_char_p
        L.soi_last_status.restype = ctypes.c_int
        L.soi_stream.argtypes = [P]
        L.soi_stream.restype = vp
        L.soi_enable_timing.argtypes = [P, ctypes.c_int]
        L.soi_stage_times.argtypes = [P, dp, ctypes.c_int]
        L.soi_stage_times.restype = ctypes.c_int
        L.soi_filter_form.argtypes = [P]
        L.soi_filter_form.restype = ctypes.c_int
        L.soi_vlc_exponents.argtypes = [P, ip, ip, ip]
        L.soi_vlc_exponents.restype = ctypes.c_int
        L.max_exponent.argtypes = [vp, ctypes.c_int]
        L.max_exponent.restype = ctypes.c_int
        L.compress.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.compress.restype = ctypes.c_int
        L.decompress.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.soi_max_exponent_batch.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.soi_compress_batch.argtypes = [vp, ctypes.c_long, vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, vp, vp, vp]
        L.soi_decompress_batch.argtypes = [vp, vp, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, vp, vp]
        L.soi_fill_uniform.argtypes = [vp, ctypes.c_long, ctypes.c_uint64, vp]
        L.soi_fill_uniform.restype = ctypes.c_int
        L.soi_fill_tones.argtypes = [vp, ctypes.c_long, dp, dp, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_uint64, vp]
        L.soi_fill_tones.restype = ctypes.c_int
        L.soi_fft_variants.argtypes = [P]
        L.soi_fft_variants.restype = ctypes.c_int
        L.soi_set_fft_variant.argtypes = [P, ctypes.c_int]
        L.soi_set_fft_variant.restype = ctypes.c_int
        L.soi_autotune.argtypes = [P, vp, ctypes.c_int]
        L.soi_autotune.restype = ctypes.c_int
        L.soi_batch_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(vp), ctypes.c_int,
                                       ctypes.c_int]
        L.soi_batch_create.restype = vp
        L.soi_batch_run.argtypes = [vp, vp]
        L.soi_batch_run.restype = ctypes.c_int
        L.soi_batch_stream.argtypes = [vp]
        L.soi_batch_stream.restype = vp
        L.soi_batch_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _check():
    L = lib()
    st = L.soi_last_status()
    if st:
        raise SoiError(st, L.soi_last_error().decode())


def set_default_soi_descriptor(d):
    lib().set_default_soi_descriptor(ctypes.byref(d))


def init_soi_descriptor(d, k=1, comm=0):
    lib().init_soi_descriptor(ctypes.byref(d), comm, k)
    _check()


def compute_soi(d, ptr):
    lib().compute_soi(ctypes.byref(d), ctypes.c_void_p(ptr))
    _check()


def compute_soi_async(d, ptr, stream=None):
    lib().compute_soi_async(ctypes.byref(d), ctypes.c_void_p(ptr), ctypes.c_void_p(stream))
    _check()


def parallel_filter_subsampling(d, ptr):
    lib().parallel_filter_subsampling(ctypes.byref(d), ctypes.c_void_p(ptr))
    _check()


def free_soi_descriptor(d):
    lib().free_soi_descriptor(ctypes.byref(d))
    _check()


def stream_of(d):
    return lib().soi_stream(ctypes.byref(d))


def enable_timing(d, on=True):
    """Start (and reset) or stop recording per-stage HIP event times of each call."""
    lib().soi_enable_timing(ctypes.byref(d), int(on))
    _check()


def stage_times(d):
    """Mean per-stage milliseconds [filter+S-DFT, vlc, fft pass 1, fft pass 2 + demodulation]
    over the recorded calls (the last 64); [] if none was recorded."""
    buf = (ctypes.c_double * 4)()
    n = lib().soi_stage_times(ctypes.byref(d), buf, 4)
    _check()
    return list(buf[:n])


# filter forms (soi_filter_form)
SOI_FILTER_VALU, SOI_FILTER_MFMA_BF16X3, SOI_FILTER_MFMA_F32, SOI_FILTER_MFMA_F64 = 0, 1, 2, 3


def filter_form(d):
    f = lib().soi_filter_form(ctypes.byref(d))
    _check()
    return f


def vlc_exponents(d):
    S = d.S
    seg = (ctypes.c_int * S)()
    tot = ctypes.c_int()
    cnt = (ctypes.c_int * S)()
    n = lib().soi_vlc_exponents(ctypes.byref(d), seg, ctypes.byref(tot), cnt)
    if n == 0:
        return None
    return list(seg), tot.value, list(cnt)


def fft_variants(d):
    n = lib().soi_fft_variants(ctypes.byref(d))
    _check()
    return n


def set_fft_variant(d, v):
    lib().soi_set_fft_variant(ctypes.byref(d), v)
    _check()


def autotune(d, scratch_ptr, reps=3):
    """Time each FFT variant on the device buffer at scratch_ptr (overwritten); keep the best."""
    v = lib().soi_autotune(ctypes.byref(d), ctypes.c_void_p(scratch_ptr), reps)
    _check()
    return v


class Batch:
    """Batching form: initialised descriptors + their device buffers, captured into one HIP
    graph (soi_batch_create); run() replays it with one launch, in place on the buffers."""

    def __init__(self, descs, ptrs, autotune=True):
        n = len(descs)
        P = ctypes.POINTER(SoiDesc)
        self._descs = descs
        arr = (P * n)(*[ctypes.pointer(d) for d in descs])
        bufs = (ctypes.c_void_p * n)(*ptrs)
        self._h = lib().soi_batch_create(arr, bufs, n, int(autotune))
        _check()
        if not self._h:
            raise SoiError(lib().soi_last_status(), lib().soi_last_error().decode())

    def run(self, stream=None):
        lib().soi_batch_run(self._h, ctypes.c_void_p(stream))
        _check()

    @property
    def stream(self):
        return lib().soi_batch_stream(self._h)

    def close(self):
        if self._h:
            lib().soi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def make_descriptor(N, P, k=1, n_mu=5, d_mu=4, B=72, tau=None, sigma=None, use_vlc=False,
                    comm_to_comp_cost_ratio=1.0, precision=SOI_PREC_FP64,
                    window_mode=SOI_WINDOW_AUTO):
    """set_default_soi_descriptor + the caller-set fields (test.c:91-101, 125-128) + init."""
    d = SoiDesc()
    set_default_soi_descriptor(d)
    d.N, d.P, d.rank = N, P, 0
    d.n_mu, d.d_mu, d.B = n_mu, d_mu, B
    if tau is not None:
        d.tau = tau
    if sigma is not None:
        d.sigma = sigma
    d.use_vlc = int(use_vlc)
    d.comm_to_comp_cost_ratio = comm_to_comp_cost_ratio
    d.precision = precision
    d.window_mode = window_mode
    init_soi_descriptor(d, k)
    return d


def soi_fft(x, P, k=1, **params):
    """Convenience: in-place SOI FFT of a CUDA(HIP) complex128 (FP64 form) or complex64 (FP32
    form) torch tensor of length N.

    Returns x (overwritten with the N-point forward DFT, natural order)."""
    import torch  # plumbing only: device memory and synchronisation
    assert x.is_cuda and x.is_contiguous() and x.dtype in (torch.complex128, torch.complex64)
    if x.dtype == torch.complex64:
        params.setdefault("precision", SOI_PREC_FP32)
    d = make_descriptor(x.numel(), P, k, **params)
    try:
        torch.cuda.synchronize()
        compute_soi(d, x.data_ptr())
    finally:
        free_soi_descriptor(d)
    return x
