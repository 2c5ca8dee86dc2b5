"""The FP32 precision form (precision = SOI_PREC_FP32): complex64 input and output, FP32 filter
contraction and FFT arithmetic, windows / twiddles generated in FP64 and W_inv applied in FP64.

Error is stated against the FP64 transform of the same complex64 input (the oracle, or rocFFT's
FP64 FFT at N = 2^27), so input rounding is not counted.  Bound: relative L2 <= 1e-6 for the
default window (max|W_inv| = 9.06); narrower windows amplify FP32 rounding near segment edges
by their demodulation gain G = max|W_inv|, bounded by 2^-24 * G."""
import ctypes

import numpy as np
import pytest

from gpu_util import rel, rel_t

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-6


def fp32_tol(orc, p):
    gain = float(np.abs(orc.w_inv_table(p)).max())
    return max(FP32_TOL, 2.0 ** -24 * gain)


def run_soi32(soi, torch, x32, P, k=1, **kw):
    d = soi.make_descriptor(x32.size, P, k, precision=soi.SOI_PREC_FP32, **kw)
    try:
        xt = torch.from_numpy(np.ascontiguousarray(x32)).cuda()
        torch.cuda.synchronize()
        soi.compute_soi(d, xt.data_ptr())
        return xt.cpu().numpy()
    finally:
        soi.free_soi_descriptor(d)


SHAPES32 = [
    (2 ** 13, 4, 1, {}),
    (2 ** 16, 8, 1, {}),
    (2 ** 16, 4, 2, {}),
    (2 ** 20, 8, 1, {}),
    (5 * 2 ** 14, 8, 1, {}),
    (2 ** 18, 32, 1, {}),
    (2 ** 17, 8, 1, dict(B=64, tau=818 / 1024, sigma=267.4513)),
    (7 * 2 ** 15, 8, 1, dict(n_mu=8, d_mu=7, tau=0.7, sigma=360.0)),
]


@pytest.mark.parametrize("N,P,k,kw", SHAPES32)
def test_fp32_form_matches_fp64_oracle(soi, orc, torch_cuda, N, P, k, kw):
    print(f"case N={N} P={P} k={k} {kw}", flush=True)
    x32 = orc.uniform_input(N, 202).astype(np.complex64)
    got = run_soi32(soi, torch_cuda, x32, P, k, **kw)
    assert got.dtype == np.complex64
    okw = {kk: v for kk, v in kw.items() if kk in ("n_mu", "d_mu", "B", "tau", "sigma")}
    p = orc.params(N, P, k=k, **okw)
    ref, _ = orc.compute_soi(p, x32.astype(np.complex128))
    err = rel(got.astype(np.complex128), ref)
    print(f"  fp32 rel-L2 vs fp64 oracle {err:.3e}", flush=True)
    assert err < fp32_tol(orc, p)


def test_fp32_host_pointer_path(soi, orc, torch_cuda):
    N, P = 2 ** 16, 8
    x32 = orc.uniform_input(N, 7).astype(np.complex64)
    d = soi.make_descriptor(N, P, precision=soi.SOI_PREC_FP32)
    try:
        buf = x32.copy()
        soi.compute_soi(d, buf.ctypes.data)
    finally:
        soi.free_soi_descriptor(d)
    ref, _ = orc.compute_soi(orc.params(N, P), x32.astype(np.complex128))
    assert rel(buf.astype(np.complex128), ref) < FP32_TOL


def test_fp32_full_size_vs_rocfft(soi, torch_cuda):
    """N = 2^27, P = 32: the tone + noise input cast to complex64, against rocFFT's FP64 FFT of
    the cast values."""
    torch = torch_cuda
    N, P = 2 ** 27, 32
    rng = np.random.default_rng(2027)
    f = np.floor(rng.uniform(0, N, 128)) + rng.uniform(0, 1, 128)
    ph = rng.uniform(0, 1, 128)
    x64 = torch.empty(N, dtype=torch.complex128, device="cuda")
    dp = ctypes.POINTER(ctypes.c_double)
    assert soi.lib().soi_fill_tones(x64.data_ptr(), N, f.ctypes.data_as(dp), ph.ctypes.data_as(dp),
                                    len(f), 1.0, 1e-3, 2027, None) == 0
    x32 = x64.to(torch.complex64)
    ref = torch.fft.fft(x32.to(torch.complex128))
    del x64
    d = soi.make_descriptor(N, P, precision=soi.SOI_PREC_FP32)
    try:
        torch.cuda.synchronize()
        soi.compute_soi(d, x32.data_ptr())
    finally:
        soi.free_soi_descriptor(d)
    err = (torch.linalg.norm(x32.to(torch.complex128) - ref) / torch.linalg.norm(ref)).item()
    print("N=2^27 fp32 rel-L2 vs rocFFT fp64", err, flush=True)
    assert err < FP32_TOL


@pytest.mark.parametrize("tiled", ["0", "1"])
def test_fp32_fft_variants_agree(soi, orc, torch_cuda, monkeypatch, tiled):
    """Every FFT variant of an FP32 plan (M_hat = 2560 x 2048 at N = 2^24, P = 4) stays within
    the FP32 bound of the FP64 oracle."""
    torch = torch_cuda
    monkeypatch.setenv("SOI_FFT_TILED", tiled)
    N, P = 2 ** 24, 4
    x32 = orc.uniform_input(N, 12).astype(np.complex64)
    ref, _ = orc.compute_soi(orc.params(N, P), x32.astype(np.complex128))
    d = soi.make_descriptor(N, P, precision=soi.SOI_PREC_FP32)
    try:
        nv = soi.fft_variants(d)
        assert nv >= 2
        for v in range(nv):
            soi.set_fft_variant(d, v)
            xt = torch.from_numpy(x32.copy()).cuda()
            torch.cuda.synchronize()
            soi.compute_soi(d, xt.data_ptr())
            err = rel(xt.cpu().numpy().astype(np.complex128), ref)
            print(f"  variant {v}: rel-L2 {err:.3e}", flush=True)
            assert err < FP32_TOL
    finally:
        soi.free_soi_descriptor(d)


MFMA_SHAPES = [(2 ** 13, 4), (2 ** 16, 8), (2 ** 18, 16), (2 ** 18, 32), (2 ** 20, 8)]


@pytest.mark.parametrize("mode", ["0", "1", "2"])
@pytest.mark.parametrize("N,P", MFMA_SHAPES)
def test_fp32_mfma_filter_matches_fp64_oracle(soi, orc, torch_cuda, monkeypatch, N, P, mode):
    """The three forms of the complex64 filter contraction for S = 4 .. 32 lanes, one and
    several tiles per block, against the FP64 oracle at the FP32 bound: SOI_FILTER_MFMA=0 on
    VALU, =1 on MFMA with every operand split into hi + mid + lo bf16 pieces (six piece
    products), =2 (the default) on MFMA with exact FP32 products (v_mfma_f32_16x16x4_f32); all
    accumulate in FP32."""
    monkeypatch.setenv("SOI_FILTER_MFMA", mode)
    print(f"case N={N} P={P} mfma filter mode {mode}", flush=True)
    x32 = orc.uniform_input(N, 303).astype(np.complex64)
    got = run_soi32(soi, torch_cuda, x32, P)
    p = orc.params(N, P)
    ref, _ = orc.compute_soi(p, x32.astype(np.complex128))
    err = rel(got.astype(np.complex128), ref)
    print(f"  mfma fp32 rel-L2 vs fp64 oracle {err:.3e}", flush=True)
    assert err < FP32_TOL


@pytest.mark.parametrize("mode", ["1", "2"])
def test_fp32_mfma_filter_forms_full_size(soi, torch_cuda, monkeypatch, mode):
    """Both MFMA forms of the complex64 filter at N = 2^27, P = 32 (the FP32-split workload):
    SOI_FILTER_MFMA=1 (operands split into bf16 pieces) and =2 (exact FP32 products, the
    default), against rocFFT's FP64 FFT of the same complex64 input at the FP32 bound."""
    torch = torch_cuda
    monkeypatch.setenv("SOI_FILTER_MFMA", mode)
    N, P = 2 ** 27, 32
    print(f"case N=2^27 P=32 mfma filter mode {mode}", flush=True)
    x = torch.empty(N, dtype=torch.complex128, device="cuda")
    assert soi.lib().soi_fill_uniform(x.data_ptr(), N, 404, None) == 0
    x32 = x.to(torch.complex64)
    del x
    ref = torch.fft.fft(x32.to(torch.complex128))
    d = soi.make_descriptor(N, P, precision=soi.SOI_PREC_FP32)
    try:
        assert soi.filter_form(d) == int(mode)
        y = x32.clone()
        torch.cuda.synchronize()
        soi.compute_soi(d, y.data_ptr())
        torch.cuda.synchronize()
        err = rel_t(torch, y.to(torch.complex128), ref)
        print(f"  N=2^27 mfma mode {mode}: rel-L2 vs rocFFT FP64 {err:.3e}", flush=True)
        assert err < FP32_TOL
    finally:
        soi.free_soi_descriptor(d)
