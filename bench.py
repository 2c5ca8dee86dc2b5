"""Benchmark: plain FP64 SOI FFT on one MI355X at N = 2^27 complex128 (2 GiB), P = 32 virtual
ranks, k = 1, mu = 5/4, B = 72; input = 128 unit-amplitude tones at random fractional bins plus
complex Gaussian noise sigma 1e-3, seeded.

One step = compute_soi over the device-resident N-vector (filter + S-DFT, segment FFT passes,
demodulation).  compute_soi works in place, so timed step i transforms one of up to 8 copies of
the input: the work of every step is identical.  Per-kernel durations come from HIP events that
the library records on its stream around each kernel during the timed steps.  Beside the
headline FP64 form the line carries the FP32 precision form at the same shape, the batching form
(512 transforms of uneven N and P, both precisions), the banded-spectrum VLC workload
(N = 7*2^23, P = 64, mu = 8/7), rocFFT on the same input in both precisions (the Cooley-Tukey
comparison) and the CPU baseline (a C/OpenMP port of the reference's timed region on this
host's cores).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-cpu-baseline] [--no-fp32]
                    [--no-batch] [--no-vlc] [--no-rocfft]
    python bench.py --copy-calibration ...   also runs one plain copy kernel (torch.neg over the
                                             N-vector: 16N bytes read, 16N written) before the
                                             timed steps, for calibrating HBM byte counters
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GFLOP/s (5N·log2N/t) at N=2^27: FP32-split vs FP64 at stated rel-L2 error"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64, vector and matrix (dense)
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X FP32 matrix (dense)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--P", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log2n", type=int, default=27)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-vlc", action="store_true")
    ap.add_argument("--no-rocfft", action="store_true")
    ap.add_argument("--batch-count", type=int, default=512)
    return ap.parse_args()


def tones(N, seed=2027, n=128):
    import numpy as np
    rng = np.random.default_rng(seed)
    f = np.floor(rng.uniform(0, N, n)) + rng.uniform(0, 1, n)
    ph = rng.uniform(0, 1, n)
    return f, ph


def traffic_from_profiles(kernel_key):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy-kernel calibration), recorded in profiles/traffic.md as 'key: bytes'."""
    path = os.path.join(ROOT, "profiles", "traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip().strip("`")
        if line.startswith(kernel_key + ":"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def fft_split(Mh):
    """n1 x n2 split of the segment FFT as the library plans it for the register kernels: both
    lengths in the supported set, n2 a multiple of 8, the larger factor as small as possible."""
    lens = [256, 320, 512, 640, 1024, 1280, 2048, 2560]
    best = None
    for n1 in lens:
        if Mh % n1:
            continue
        n2 = Mh // n1
        if n2 in lens[:7] and n2 % 8 == 0 and (best is None or max(n1, n2) < max(best)):
            best = (n1, n2)
    return best if best else (Mh, 1)


def errors(torch, y, ref):
    """Relative L2, max-abs and the normalised infinity norm max|y - ref| / max|ref|
    (input.c:395-425) of y against ref, both on the device."""
    d = (y - ref)
    rel = (torch.linalg.norm(d) / torch.linalg.norm(ref)).item()
    mx = d.abs().max().item()
    return {"rel_l2": rel, "max_abs": mx, "normalized_inf_norm": mx / ref.abs().max().item()}


def step_buffers(torch, x, steps):
    """Up to 8 device copies of the input; timed step i transforms copy i % 8 in place, so a
    copy is transformed at most ceil(steps / 8) times (each transform scales the norm by
    sqrt(N), which a complex64 copy tolerates a few times over before overflowing)."""
    return [x.clone() for _ in range(max(1, min(steps, 8)))]


def refill(torch, bufs, x):
    for b in bufs:
        b.copy_(x)


def timed_steps(args, soi, torch, d, bufs, x):
    """warmup, refill, then args.steps timed compute_soi calls on the descriptor's stream"""
    stream = torch.cuda.ExternalStream(soi.stream_of(d))
    for _ in range(args.warmup):
        soi.compute_soi_async(d, bufs[0].data_ptr(), None)
    refill(torch, bufs, x)
    torch.cuda.synchronize()
    soi.enable_timing(d, True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(args.steps):
        soi.compute_soi_async(d, bufs[i % len(bufs)].data_ptr(), None)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    acc = soi.stage_times(d)
    soi.enable_timing(d, False)
    return ms, acc, wall


def kernel_table(N, S, M, Mh, mu, B, acc, esz):
    """algorithmic bytes and flops per launch of each kernel (one launch covers all S segments);
    esz = bytes per complex value"""
    n1, n2 = fft_split(Mh)
    return {
        "k_filter": (acc[0], esz * N * (1 + mu), (4.0 * B + 6 + 5 * math.log2(S)) * mu * N),
        "k_fft_cols": (acc[2], 2.0 * esz * mu * N, (5.0 * math.log2(n1) + 6) * mu * N),
        "k_fft_rows": (acc[3], esz * mu * N + esz * N + 16.0 * M,
                       5.0 * math.log2(n2) * mu * N + 6.0 * N),
    }, (n1, n2)


def roofline_of(kernels, filter_form, soi, traffic_key):
    """Roofline of the dominant (longest) kernel: HBM for the FFT passes; for the filter the
    unit its contraction runs on (MFMA or the FP64 vector units) or HBM, whichever it is
    closer to."""
    name = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes, kflops = kernels[name]
    gbs = kbytes / (kms * 1e-3) / 1e9
    tfs = kflops / (kms * 1e-3) / 1e12
    hbm_frac = gbs / HBM_PEAK_GBS
    if name == "k_filter" and filter_form == soi.SOI_FILTER_MFMA_F32:
        cpeak, cbound = FP32_MFMA_PEAK_TFLOPS, "mfma"
    elif name == "k_filter" and filter_form == soi.SOI_FILTER_MFMA_F64:
        cpeak, cbound = FP64_PEAK_TFLOPS, "mfma"
    else:
        cpeak, cbound = FP64_PEAK_TFLOPS, "fp64"   # VALU FMAs
    c_frac = tfs / cpeak
    hbm = name != "k_filter" or hbm_frac >= c_frac
    return {"bound": "hbm" if hbm else cbound,
            "achieved": round(gbs if hbm else tfs, 2),
            "peak": HBM_PEAK_GBS if hbm else cpeak,
            "unit": "GB/s" if hbm else "TFLOP/s",
            "frac": round(hbm_frac if hbm else c_frac, 4),
            "traffic": traffic_from_profiles(f"{name} {traffic_key}"),
            "kernel": name, "launch_ms": round(kms, 4),
            "algorithmic_bytes": kbytes, "algorithmic_flops": kflops,
            "hbm_frac": round(hbm_frac, 4), "compute_frac": round(c_frac, 4),
            "compute_unit": "v_mfma_f64_16x16x4_f64" if cbound == "mfma" and cpeak == FP64_PEAK_TFLOPS
            else "v_mfma_f32_16x16x4_f32" if cbound == "mfma" else "FP64 VALU"}


def bench_fp32(args, soi, torch, x64, ref, N, P, fp64_gflops):
    """The FP32 precision form on the same input rounded to complex64: throughput over the same
    number of timed steps; errors against rocFFT's FP64 FFT of the rounded input (input
    rounding is not counted)."""
    x = x64.to(torch.complex64)
    d = soi.make_descriptor(N, P, precision=soi.SOI_PREC_FP32)
    form = soi.filter_form(d)
    ref32 = torch.fft.fft(x.to(torch.complex128))
    y = x.clone()
    torch.cuda.synchronize()
    soi.compute_soi(d, y.data_ptr())
    err = errors(torch, y.to(torch.complex128), ref32)
    del ref32, y
    torch.cuda.empty_cache()
    bufs = step_buffers(torch, x, args.steps)
    ms, acc, _ = timed_steps(args, soi, torch, d, bufs, x)
    del bufs
    S, M, Mh = d.S, d.M, d.M_hat
    soi.free_soi_descriptor(d)
    kernels, _ = kernel_table(N, S, M, Mh, 5 / 4, 72, acc, 8)
    gflops = 5.0 * N * math.log2(N) / (ms * 1e-3) / 1e9
    return {"value": round(gflops, 2), "unit": "GFLOP/s", "ms_per_step": round(ms, 4),
            "dtype": "f32", "errors_vs_rocfft_fp64": err,
            "rel_l2_vs_rocfft_fp64": err["rel_l2"],
            "speedup_vs_fp64_form": round(gflops / fp64_gflops, 3),
            "stages": {"filter_sdft_ms": round(acc[0], 4), "fft_pass1_ms": round(acc[2], 4),
                       "fft_pass2_demod_ms": round(acc[3], 4)},
            "roofline": roofline_of(kernels, form, soi, f"fp32 N=2^{args.log2n} P={P}"),
            "form": "complex64 data; FP32 filter contraction on MFMA (v_mfma_f32_16x16x4_f32, "
                    "exact FP32 products) and FP32 FFTs; FP64 window, twiddle and W_inv "
                    "generation; W_inv applied in FP64"}


def bench_fp32_bf16x3(args, soi, torch, x64, N, P):
    """The FP32 form with the split-bf16 MFMA filter (SOI_FILTER_MFMA=1: every operand split
    into hi + mid + lo bf16 pieces, six piece products on v_mfma_f32_16x16x32_bf16), opt-in, at
    the same shape: its filter time and errors beside the default exact-FP32 MFMA filter."""
    old = os.environ.get("SOI_FILTER_MFMA")
    os.environ["SOI_FILTER_MFMA"] = "1"
    try:
        d = soi.make_descriptor(N, P, precision=soi.SOI_PREC_FP32)
    finally:
        if old is None:
            del os.environ["SOI_FILTER_MFMA"]
        else:
            os.environ["SOI_FILTER_MFMA"] = old
    form = soi.filter_form(d)
    x = x64.to(torch.complex64)
    ref32 = torch.fft.fft(x.to(torch.complex128))
    y = x.clone()
    torch.cuda.synchronize()
    soi.compute_soi(d, y.data_ptr())
    err = errors(torch, y.to(torch.complex128), ref32)
    del ref32, y
    torch.cuda.empty_cache()
    a = argparse.Namespace(steps=max(3, args.steps // 2), warmup=min(args.warmup, 3))
    bufs = step_buffers(torch, x, a.steps)
    ms, acc, _ = timed_steps(a, soi, torch, d, bufs, x)
    del bufs
    soi.free_soi_descriptor(d)
    return {"filter_form": form, "ms_per_step": round(ms, 4), "steps": a.steps,
            "filter_sdft_ms": round(acc[0], 4), "errors_vs_rocfft_fp64": err}


def bench_batch(args, soi, torch, prec):
    """The batching form: args.batch_count independent transforms, N uniform over {2^16, 3*2^15,
    2^18, 5*2^16, 2^20}, P uniform over {8, 16, 32} (rejecting (2^16, 32), whose M/P < B),
    uniform[-1,1] inputs, captured into one HIP graph (each shape autotuned once) and replayed:
    one step = one graph launch transforming all of them.  Every buffer is refilled from its
    pristine copy before each replay, outside the timed region (each replay has its own event
    pair); afterwards one member of each distinct shape is compared with an eager compute_soi
    of its input."""
    import numpy as np
    sizes, ps = [2 ** 16, 3 * 2 ** 15, 2 ** 18, 5 * 2 ** 16, 2 ** 20], [8, 16, 32]
    rng = np.random.default_rng(2028)
    shapes = []
    while len(shapes) < args.batch_count:
        n, p = int(rng.choice(sizes)), int(rng.choice(ps))
        if (n, p) != (2 ** 16, 32):
            shapes.append((n, p))
    cdt = torch.complex64 if prec else torch.complex128
    descs, bufs, pristine = [], [], []
    t0 = time.perf_counter()
    for i, (n, p) in enumerate(shapes):
        descs.append(soi.make_descriptor(n, p, precision=prec))
        b = torch.empty(n, dtype=torch.complex128, device="cuda")
        soi.lib().soi_fill_uniform(b.data_ptr(), n, 7000 + i, None)
        b = b.to(cdt) if prec else b
        pristine.append(b.clone())
        bufs.append(b)
    torch.cuda.synchronize()
    batch = soi.Batch(descs, [b.data_ptr() for b in bufs], autotune=True)
    setup_s = time.perf_counter() - t0
    stream = torch.cuda.ExternalStream(batch.stream)

    def refill_all():
        for b, q in zip(bufs, pristine):
            b.copy_(q)
        torch.cuda.synchronize()

    for _ in range(args.warmup):
        refill_all()
        batch.run()
    total = 0.0
    for _ in range(args.steps):
        refill_all()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        batch.run()
        e1.record(stream)
        torch.cuda.synchronize()
        total += e0.elapsed_time(e1)
    ms = total / args.steps
    # the last replay transformed the pristine inputs: check one member per distinct shape
    checked, worst, tol = {}, 0.0, 1e-6 if prec else 1e-12
    for i, sh in enumerate(shapes):
        if sh in checked:
            continue
        y = pristine[i].clone()
        d = soi.make_descriptor(sh[0], sh[1], precision=prec)
        torch.cuda.synchronize()
        soi.compute_soi(d, y.data_ptr())
        soi.free_soi_descriptor(d)
        out = bufs[i]
        assert bool(torch.isfinite(torch.view_as_real(out)).all()), f"member {i} not finite"
        e = (torch.linalg.norm((out - y).to(torch.complex128)) /
             torch.linalg.norm(y.to(torch.complex128))).item()
        checked[sh] = e
        worst = max(worst, e)
    assert worst <= tol, f"batch member differs from eager compute_soi: {worst}"
    flop = sum(5.0 * n * math.log2(n) for n, _ in shapes)
    batch.close()
    for d in descs:
        soi.free_soi_descriptor(d)
    del bufs, pristine
    torch.cuda.empty_cache()
    return {"value": round(flop / (ms * 1e-3) / 1e9, 2), "unit": "GFLOP/s",
            "ms_per_step": round(ms, 4), "transforms": len(shapes),
            "points": int(sum(n for n, _ in shapes)), "setup_s": round(setup_s, 2),
            "dtype": "f32" if prec else "f64",
            "max_rel_l2_vs_eager": worst, "shapes_checked": len(checked),
            "launch": "one HIP graph launch per step; inside it one launch per kernel per group of "
                      "same-shape transforms (blockIdx.y = transform), groups over 4 capture branches"}


def bench_vlc(args, soi, torch):
    """The banded-spectrum workload with VLC compression of the exchanged data: N = 7*2^23
    complex128, P = 64, mu = 8/7 (tau = 0.7, sigma = 360), B = 72, use_vlc = 1; input = 256
    unit tones inside one band of width N/64 plus white noise at -60 dB relative to the tone
    power, seeded."""
    import numpy as np
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
    d = soi.make_descriptor(N, P, use_vlc=True, **kw)
    y = x.clone()
    torch.cuda.synchronize()
    soi.compute_soi(d, y.data_ptr())
    err = errors(torch, y, ref)
    seg, tot, cnt = soi.vlc_exponents(d)
    del ref, y
    torch.cuda.empty_cache()
    steps = max(3, args.steps // 2)
    bufs = [x.clone() for _ in range(2)]
    a = argparse.Namespace(steps=steps, warmup=min(args.warmup, 3))
    ms, acc, _ = timed_steps(a, soi, torch, d, bufs, x)
    soi.free_soi_descriptor(d)
    del bufs, x
    torch.cuda.empty_cache()
    full = 2 * 8 * N // 7  # doubles of alpha_tilde: S * M_hat complex values
    return {"value": round(5.0 * N * math.log2(N) / (ms * 1e-3) / 1e9, 2), "unit": "GFLOP/s",
            "ms_per_step": round(ms, 4), "steps": steps, "dtype": "f64",
            "config": {"workload": "banded-spectrum SOI with VLC, N=7*2^23 complex128, P=64, "
                                   "mu=8/7 (tau=0.7, sigma=360), B=72, use_vlc=1",
                       "N": N, "P": P, "n_mu": 8, "d_mu": 7, "B": 72},
            "stages": {"filter_sdft_ms": round(acc[0], 4), "vlc_ms": round(acc[1], 4),
                       "fft_pass1_ms": round(acc[2], 4), "fft_pass2_demod_ms": round(acc[3], 4)},
            "errors_vs_rocfft_fp64": err,
            "vlc_send_doubles": int(sum(cnt)), "alpha_tilde_doubles": full,
            "vlc_total_max_exponent": tot,
            "note": "N=2^26 has no valid (P*d_mu) | M split at mu=8/7; 7*2^23 is the nearest"}


def bench_rocfft(args, torch, x, soi_ms):
    """The Cooley-Tukey comparison (the role of the reference test.c's MKL / MKL-CDFT leg,
    test.c:246-271): rocFFT through torch.fft.fft (hipFFT) on the same N-point input, out of
    place into a preallocated output, complex128 and complex64, timed with events like the SOI
    steps.  soi_speedup = rocFFT ms / SOI ms of the same precision."""
    out = {}
    N = x.numel()
    flop = 5.0 * N * math.log2(N)
    for name, dt in (("fp64", torch.complex128), ("fp32", torch.complex64)):
        if name not in soi_ms:
            continue
        xi = x if dt == torch.complex128 else x.to(dt)
        y = torch.empty_like(xi)
        for _ in range(max(2, args.warmup)):
            torch.fft.fft(xi, out=y)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            torch.fft.fft(xi, out=y)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        out[name] = {"ms_per_step": round(ms, 4), "GFLOP/s": round(flop / (ms * 1e-3) / 1e9, 2),
                     "soi_speedup": round(ms / soi_ms[name], 3)}
        del y, xi
        torch.cuda.empty_cache()
    out["note"] = "torch.fft.fft (hipFFT / rocFFT) of the same input, out of place"
    return out


def cpu_baseline(log2n, P):
    """The CPU performance port of the reference's timed region (test.c:389-404: compute_soi on
    a resident input, tables built beforehand; oracle/soi_cpu_port.c, C + OpenMP + AVX2) on this
    host's cores: N = 2^log2n with the same S, mu, B as the GPU workload, best of two runs."""
    import numpy as np
    import oracle
    N = 1 << log2n
    p = oracle.params(N, P)
    x = oracle.uniform_input(N, 5)
    t = time.perf_counter()
    port = oracle.CpuPort(p)
    setup = time.perf_counter() - t
    best = None
    for _ in range(2):
        y = x.copy()
        t = time.perf_counter()
        port.compute(y)
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    port.close()
    return {"value": 5 * N * math.log2(N) / best / 1e9, "unit": "GFLOP/s",
            "cores": int(oracle.port_threads()), "kind": "port",
            "sample": f"CPU port of test.c:389-404's timed region (oracle/soi_cpu_port.c: "
                      f"complex-tap conv with AVX2 FMA, four-step Stockham FFTs, fused "
                      f"demodulation), N=2^{log2n} complex128, P={P}, k=1, mu=5/4, B=72, "
                      f"uniform input; best of 2 runs: {best:.3f} s per transform "
                      f"(plan {setup:.2f} s, not timed)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU "
                                   "sharding is a design only)"}))
        sys.exit(2)
    import numpy as np
    import torch
    import soi_fft_amd as soi

    N, P = 1 << args.log2n, args.P
    dev = torch.device("cuda:0")
    f, ph = tones(N)
    x = torch.empty(N, dtype=torch.complex128, device=dev)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = soi.lib().soi_fill_tones(x.data_ptr(), N, f.ctypes.data_as(dp), ph.ctypes.data_as(dp),
                                  len(f), 1.0, 1e-3, 2027, None)
    assert rc == 0

    d = soi.make_descriptor(N, P)
    form = soi.filter_form(d)
    # accuracy of the step as timed: one transform against rocFFT's FP64 FFT
    ref = torch.fft.fft(x)
    y = x.clone()
    torch.cuda.synchronize()
    soi.compute_soi(d, y.data_ptr())
    err = errors(torch, y, ref)
    if args.copy_calibration:
        torch.neg(x, out=y)
        torch.cuda.synchronize()
    del ref, y
    torch.cuda.empty_cache()

    bufs = step_buffers(torch, x, args.steps)
    ms, acc, wall = timed_steps(args, soi, torch, d, bufs, x)
    del bufs
    flop = 5.0 * N * math.log2(N)
    gflops = flop / (ms * 1e-3) / 1e9
    S, M, Mh = d.S, d.M, d.M_hat
    soi.free_soi_descriptor(d)

    mu, B = 5 / 4, 72
    kernels, (n1, n2) = kernel_table(N, S, M, Mh, mu, B, acc, 16)
    roofline = roofline_of(kernels, form, soi, f"N=2^{args.log2n} P={P}")
    stages = {"filter_sdft_ms": round(acc[0], 4), "vlc_ms": round(acc[1], 4),
              "fft_pass1_ms": round(acc[2], 4), "fft_pass2_demod_ms": round(acc[3], 4)}
    whole_bytes = 16.0 * N * (2 + 2 * mu)
    out = {
        "metric": METRIC, "value": round(gflops, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: 128 unit-amplitude tones at random fractional bins + complex "
                "Gaussian noise sigma=1e-3 per component, seeded",
        "config": {"workload": f"plain SOI FP64, N=2^{args.log2n} complex128, P={P} virtual "
                               "ranks, k=1, mu=5/4, B=72",
                   "N": N, "P": P, "k": 1, "S": S, "M": M, "M_hat": Mh, "n_mu": 5, "d_mu": 4,
                   "B": B, "form": "fp64", "fft_split": [n1, n2],
                   "filter": {0: "FP64 VALU", 3: "v_mfma_f64_16x16x4_f64"}.get(form, str(form))},
        "errors_vs_rocfft_fp64": err,
        "rel_l2_vs_rocfft_fp64": err["rel_l2"],
        "wall_s_timed": round(wall, 4),
        "whole_transform_hbm_roofline_frac": round(whole_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "stages": stages,
        "kernel_rooflines": {k: {"ms": round(v[0], 4),
                                 "GB/s": round(v[1] / (v[0] * 1e-3) / 1e9, 1),
                                 "TFLOP/s": round(v[2] / (v[0] * 1e-3) / 1e12, 2)}
                             for k, v in kernels.items()},
        "roofline": roofline,
    }
    if not args.no_fp32:
        out["fp32_split"] = bench_fp32(args, soi, torch, x, None, N, P, gflops)
        out["fp32_split"]["bf16x3_filter"] = bench_fp32_bf16x3(args, soi, torch, x, N, P)
    if not args.no_rocfft:
        soi_ms = {"fp64": ms}
        if "fp32_split" in out:
            soi_ms["fp32"] = out["fp32_split"]["ms_per_step"]
        out["rocfft"] = bench_rocfft(args, torch, x, soi_ms)
    del x
    torch.cuda.empty_cache()
    if not args.no_batch:
        out["batch_form"] = {"fp64": bench_batch(args, soi, torch, 0),
                             "fp32": bench_batch(args, soi, torch, 1)}
    if not args.no_vlc:
        out["vlc_banded"] = bench_vlc(args, soi, torch)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_log2n, P)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
