"""The VLC codec on the GPU is bit-exact with the oracle's restatement of compress.c, through
the reference's own entry points (compress.h) and the batched device forms; compute_soi with
use_vlc matches the oracle's VLC branch (pfft.c:316-353, 390-440, 476-543, 618-632)."""
import ctypes

import numpy as np
import pytest

from gpu_util import rel, to_dev

pytestmark = pytest.mark.gpu


def _arr(rng, n, e):
    x = rng.uniform(-1, 1, n) * 2.0 ** (e + 1) * 0.999
    return x


def test_max_exponent_host_and_device(soi, orc, torch_cuda):
    rng = np.random.default_rng(0)
    L = soi.lib()
    cases = [np.array([1.0]), np.zeros(40), np.array([5e-324, 0.0, -0.0]),
             rng.normal(size=1001) * 1e-3, rng.normal(size=16) * 1e300, np.array([np.inf, 1.0])]
    for x in cases:
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert L.max_exponent(x.ctypes.data, len(x)) == orc.max_exponent(x)
        xt = to_dev(torch_cuda, x)
        assert L.max_exponent(xt.data_ptr(), len(x)) == orc.max_exponent(x)


CASES = [(16, 0), (32, 0), (35, 0), (64, 5), (100, 23), (48, 24), (21, 30), (160, 40), (77, 54),
         (8, 3), (1, 0), (4096, 1), (32768, 0), (1000, 12)]


@pytest.mark.parametrize("length,delta", CASES)
def test_compress_bit_exact(soi, orc, torch_cuda, length, delta):
    rng = np.random.default_rng(length * 100 + delta)
    e_max = 4
    e_i = e_max - delta
    x = _arr(rng, length, e_i)
    x[rng.integers(0, length)] = 0.0
    ref = orc.compress(x, e_max, e_i)
    L = soi.lib()
    out = np.zeros(2 * length + 64, dtype=np.int32)
    n = L.compress(out.ctypes.data, x.ctypes.data, length, e_max, e_i, 0)
    assert n == len(ref)
    assert np.array_equal(out[:n], ref)
    # decompress: bit-exact doubles
    back = np.zeros(length)
    words = np.zeros(2 * length + 64, dtype=np.int32)
    words[:n] = out[:n]
    L.decompress(back.ctypes.data, words.ctypes.data, length, e_max, e_i, None)
    oref = orc.decompress(ref, length, e_max, e_i)
    assert np.array_equal(back.view(np.int64), oref.view(np.int64))


def test_compress_rounding_ties(soi, orc, torch_cuda):
    # exact ties of x*2^(21-e_max) in the vector body (half-even) and the tail (half-away)
    e_max = 0
    x = np.array([(k + 0.5) / 2.0 ** 21 for k in range(-20, 20)] + [2.5 / 2.0 ** 21] * 3)
    ref = orc.compress(x, e_max, e_max)
    out = np.zeros(256, dtype=np.int32)
    n = soi.lib().compress(out.ctypes.data, x.ctypes.data, len(x), e_max, e_max, 0)
    assert np.array_equal(out[:n], ref)


def test_batched_device_codec(soi, orc, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    nch, length = 12, 2 * 1000
    emax = 3
    emi = np.array([3, 3, 2, 0, -5, -20, -21, -30, -51, -52, -60, 3], dtype=np.int32)
    x = np.concatenate([_arr(rng, length, e) for e in emi])
    stride = length * 2 + 64
    xt = to_dev(torch, x)
    emt = to_dev(torch, emi)
    wt = torch.zeros(nch * stride, dtype=torch.int32, device="cuda")
    nwt = torch.zeros(nch, dtype=torch.int32, device="cuda")
    L = soi.lib()
    assert L.soi_compress_batch(wt.data_ptr(), stride, xt.data_ptr(), length, nch, emax,
                                emt.data_ptr(), nwt.data_ptr(), None) == 0
    yt = torch.zeros(nch * length, dtype=torch.float64, device="cuda")
    assert L.soi_decompress_batch(yt.data_ptr(), wt.data_ptr(), stride, length, nch, emax,
                                  emt.data_ptr(), None) == 0
    torch.cuda.synchronize()
    w = wt.cpu().numpy().reshape(nch, stride)
    nw = nwt.cpu().numpy()
    y = yt.cpu().numpy().reshape(nch, length)
    for c in range(nch):
        chunk = x[c * length:(c + 1) * length]
        ref = orc.compress(chunk, emax, int(emi[c]))
        assert nw[c] == len(ref)
        assert np.array_equal(w[c, :nw[c]], ref)
        assert np.array_equal(y[c], orc.decompress(ref, length, emax, int(emi[c])))
    ex = torch.zeros(nch, dtype=torch.int32, device="cuda")
    assert L.soi_max_exponent_batch(xt.data_ptr(), length, nch, ex.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert list(ex.cpu().numpy()) == [orc.max_exponent(x[c * length:(c + 1) * length]) for c in range(nch)]


def test_flat_codec_matches_batched(soi, orc, torch_cuda):
    """compute_soi's VLC exchange runs the planning kernel, then compress and decompress as
    flat grids over all chunks (per-lane runs of words).  Their words, word counts and decoded
    doubles must equal the per-chunk batched codec's bit for bit, chunk by chunk, with field
    widths from 56 bits down to none (a segment 60 binades below the largest) and a tail of 8
    elements per chunk."""
    torch = torch_cuda
    L = soi.lib()
    c_long, c_int, c_void, c_dbl = ctypes.c_long, ctypes.c_int, ctypes.c_void_p, ctypes.c_double
    L.soik_max_exponent.argtypes = [c_void, c_int, c_int, c_void, c_void]
    L.soik_vlc_plan.argtypes = [c_void, c_int, c_int, c_int, c_long, c_long, c_dbl, c_void,
                                c_void, c_void, c_void, c_void, c_void]
    L.soik_compress_flat.argtypes = [c_void, c_long, c_void, c_int, c_int, c_void, c_void,
                                     c_void, c_void]
    L.soik_decompress_flat.argtypes = [c_void, c_void, c_long, c_int, c_int, c_void, c_void,
                                       c_void]
    S, P, l = 8, 4, 1004
    k, length, nch, stride = S // P, 2 * l, S * P, 4 * l
    rng = np.random.default_rng(11)
    seg_e = [3, 2, 0, -7, -20, -31, -45, -60]
    x = np.concatenate([_arr(rng, length, seg_e[c // P] - int(rng.integers(0, 3)))
                        for c in range(nch)])
    print(f"case flat codec S={S} P={P} l={l}", flush=True)
    xt = to_dev(torch, x)
    i32 = dict(dtype=torch.int32, device="cuda")
    exps, emi = torch.zeros(nch, **i32), torch.zeros(nch, **i32)
    plan = torch.zeros(1 + 2 * S + P + 1, **i32)
    nw_flat, nw_ref = torch.zeros(nch, **i32), torch.zeros(nch, **i32)
    pcs = torch.zeros(S + 1, dtype=torch.float64, device="cuda")
    upre = torch.zeros(nch + 1, dtype=torch.int64, device="cuda")
    w_flat = torch.zeros(nch * stride, **i32)
    w_ref = torch.zeros(nch * stride, **i32)
    assert L.soik_max_exponent(xt.data_ptr(), length, nch, exps.data_ptr(), None) == 0
    assert L.soik_vlc_plan(exps.data_ptr(), S, P, k, l, P * l, 1.0, emi.data_ptr(),
                           plan.data_ptr(), pcs.data_ptr(), upre.data_ptr(),
                           nw_flat.data_ptr(), None) == 0
    assert L.soik_compress_flat(w_flat.data_ptr(), stride, xt.data_ptr(), length, nch,
                                plan.data_ptr(), emi.data_ptr(), upre.data_ptr(), None) == 0
    torch.cuda.synchronize()
    e_max = int(plan[0].item())
    assert e_max == orc.max_exponent(x)
    assert L.soi_compress_batch(w_ref.data_ptr(), stride, xt.data_ptr(), length, nch, e_max,
                                emi.data_ptr(), nw_ref.data_ptr(), None) == 0
    y_flat = torch.zeros(nch * length, dtype=torch.float64, device="cuda")
    y_ref = torch.zeros(nch * length, dtype=torch.float64, device="cuda")
    assert L.soik_decompress_flat(y_flat.data_ptr(), w_flat.data_ptr(), stride, length, nch,
                                  plan.data_ptr(), emi.data_ptr(), None) == 0
    assert L.soi_decompress_batch(y_ref.data_ptr(), w_ref.data_ptr(), stride, length, nch, e_max,
                                  emi.data_ptr(), None) == 0
    torch.cuda.synchronize()
    nwf, nwr = nw_flat.cpu().numpy(), nw_ref.cpu().numpy()
    assert np.array_equal(nwf, nwr)
    assert nwr[0] > 0 and nwr[-1] == 0  # widest fields in segment 0, none in the last
    wf = w_flat.cpu().numpy().reshape(nch, stride)
    wr = w_ref.cpu().numpy().reshape(nch, stride)
    for c in range(nch):
        assert np.array_equal(wf[c, :nwr[c]], wr[c, :nwr[c]]), c
    assert torch.equal(y_flat.view(torch.int64), y_ref.view(torch.int64))


def _banded(N, S, seed, ntones=256):
    """Tones inside one band of width N/S plus white noise at -60 dB (the banded VLC workload)."""
    rng = np.random.default_rng(seed)
    lo = rng.integers(0, S) * (N // S)
    f = lo + rng.uniform(0, N // S, ntones)
    n = np.arange(N)
    x = np.zeros(N, dtype=np.complex128)
    for fj, pj in zip(f, rng.uniform(0, 1, ntones)):
        x += np.exp(2j * np.pi * (fj * (n / N) + pj))
    p_sig = np.mean(np.abs(x) ** 2)
    s = np.sqrt(p_sig * 1e-6 / 2)
    x += s * (rng.normal(size=N) + 1j * rng.normal(size=N))
    return x


@pytest.mark.parametrize("N,P,kw", [
    (2 ** 16, 8, {}),
    (7 * 2 ** 15, 16, dict(n_mu=8, d_mu=7, tau=0.7, sigma=360.0)),
    (7 * 2 ** 17, 64, dict(n_mu=8, d_mu=7, tau=0.7, sigma=360.0)),   # banded VLC workload at N/64
])
def test_soi_vlc_matches_oracle(soi, orc, torch_cuda, N, P, kw):
    torch = torch_cuda
    x = _banded(N, P, 11)
    d = soi.make_descriptor(N, P, use_vlc=True, **kw)
    try:
        xt = to_dev(torch, x)
        soi.compute_soi(d, xt.data_ptr())
        got = xt.cpu().numpy()
        sb = list(d.segmentBoundaries[: P + 1])
        seg, tot, cnt = soi.vlc_exponents(d)
    finally:
        soi.free_soi_descriptor(d)
    p = orc.params(N, P, **kw)
    ref, osb = orc.compute_soi(p, x, use_vlc=True)
    gain = float(np.abs(orc.w_inv_table(p)).max())
    assert rel(got, ref) < max(1e-12, 2 * 2.0 ** -52 * gain)
    assert sb == list(osb)
    # per-segment exponents from the oracle's Step-1 output
    at = orc.filter_subsample(p, x)
    oseg = [orc.max_exponent(at[s].view(np.float64)) for s in range(P)]
    assert seg == oseg and tot == max(oseg)
    sc, _ = orc.vlc_plan(P, 1, at.shape[1] // P, at.shape[1], 1.0, oseg)
    assert cnt == list(sc)


def test_soi_vlc_banded_full_size(soi, orc, torch_cuda):
    """The banded VLC workload at full size: N = 7*2^23 complex128, P = 64, mu = 8/7, B = 72;
    256 unit tones inside one band of width N/64 plus white noise at -60 dB.  The VLC run must
    agree with the uncompressed run to the FP64 parity bound (the codec keeps 55 - Delta bits
    per value against the global maximum exponent), and both with rocFFT's FP64 FFT to the
    window's accuracy."""
    torch = torch_cuda
    N, P = 7 * 2 ** 23, 64
    kw = dict(n_mu=8, d_mu=7, tau=0.7, sigma=360.0)
    rng = np.random.default_rng(5)
    lo = int(rng.integers(0, P)) * (N // P)
    f = lo + rng.uniform(0, N // P, 256)
    ph = rng.uniform(0, 1, 256)
    sigma = float(np.sqrt(256 * 1e-6 / 2))
    x = torch.empty(N, dtype=torch.complex128, device="cuda")
    dp = ctypes.POINTER(ctypes.c_double)
    assert soi.lib().soi_fill_tones(x.data_ptr(), N, f.ctypes.data_as(dp), ph.ctypes.data_as(dp),
                                    len(f), 1.0, sigma, 5, None) == 0
    ref = torch.fft.fft(x)
    outs = {}
    for vlc in (False, True):
        y = x.clone()
        d = soi.make_descriptor(N, P, use_vlc=vlc, **kw)
        try:
            torch.cuda.synchronize()
            soi.compute_soi(d, y.data_ptr())
            if vlc:
                sb = list(d.segmentBoundaries[: P + 1])
                seg, tot, cnt = soi.vlc_exponents(d)
        finally:
            soi.free_soi_descriptor(d)
        outs[vlc] = y
    e0 = (torch.linalg.norm(outs[False] - ref) / torch.linalg.norm(ref)).item()
    ev = (torch.linalg.norm(outs[True] - outs[False]) / torch.linalg.norm(outs[False])).item()
    print(f"banded N=7*2^23: rel-L2 vs rocFFT {e0:.3e}, vlc vs plain {ev:.3e}", flush=True)
    gain = float(np.abs(orc.w_inv_table(orc.params(7 * 2 ** 13, 4, **kw))).max())
    assert e0 < 1e-10
    assert ev < max(1e-12, 2 * 2.0 ** -52 * gain)
    # segment boundaries: a partition of the S segments over the P ranks
    assert sb[0] == 0 and sb[-1] == P and all(b <= c for b, c in zip(sb, sb[1:]))
    assert tot == max(seg) and len(cnt) == P


@pytest.mark.parametrize("P,k,ratio", [(8, 1, 0.25), (8, 2, 3.7), (4, 4, 40.0)])
def test_soi_vlc_async_plan_matches_oracle(soi, orc, torch_cuda, P, k, ratio):
    """compute_soi_async with use_vlc plans on the device (no host round trip): after the
    stream completes, the output, the cost-balanced segmentBoundaries (pfft.c:390-440, for
    several cost ratios and k > 1) and the send counts match the oracle exactly."""
    torch = torch_cuda
    N = 2 ** 16
    x = _banded(N, P * k, 13)
    d = soi.make_descriptor(N, P, k, use_vlc=True, comm_to_comp_cost_ratio=ratio)
    try:
        xt = to_dev(torch, x)
        torch.cuda.synchronize()
        soi.compute_soi_async(d, xt.data_ptr(), None)
        torch.cuda.synchronize()
        got = xt.cpu().numpy()
        sb = list(d.segmentBoundaries[: P + 1])
        seg, tot, cnt = soi.vlc_exponents(d)
    finally:
        soi.free_soi_descriptor(d)
    p = orc.params(N, P, k)
    ref, osb = orc.compute_soi(p, x, use_vlc=True, comm_to_comp=ratio)
    assert rel(got, ref) < 1e-12
    assert sb == list(osb)
    at = orc.filter_subsample(p, x)
    oseg = [orc.max_exponent(at[s].view(np.float64)) for s in range(P * k)]
    assert seg == oseg and tot == max(oseg)
    sc, _ = orc.vlc_plan(P, k, at.shape[1] // P, at.shape[1], ratio, oseg)
    assert cnt == list(sc)
