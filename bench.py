"""Benchmark: plain FP64 SOI FFT on one MI355X at N = 2^27 complex128 (2 GiB), P = 32 virtual
ranks, k = 1, mu = 5/4, B = 72; input = 128 unit-amplitude tones at random fractional bins plus
complex Gaussian noise sigma 1e-3, seeded.

One step = compute_soi over the device-resident N-vector (filter + S-DFT, segment FFT passes,
demodulation).  compute_soi works in place, so step k transforms step k-1's output: the work of
every step is identical.  Per-kernel durations come from HIP events that the library records on
its stream around each kernel during the timed steps.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-cpu-baseline]
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
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--P", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log2n", type=int, default=25)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
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


def step_buffers(torch, x, steps):
    """Up to 8 device copies of the input; timed step i transforms copy i % 8 in place, so a
    copy is transformed at most ceil(steps / 8) times (each transform scales the norm by
    sqrt(N), which a complex64 copy tolerates a few times over before overflowing)."""
    return [x.clone() for _ in range(max(1, min(steps, 8)))]


def refill(torch, bufs, x):
    for b in bufs:
        b.copy_(x)


def bench_fp32(args, soi, torch, x64, N, P, fp64_gflops):
    """The FP32 precision form on the same input rounded to complex64: throughput over the same
    number of timed steps, and relative L2 error against rocFFT's FP64 FFT of the rounded input
    (input rounding is not counted)."""
    x = x64.to(torch.complex64)
    d = soi.make_descriptor(N, P, precision=soi.SOI_PREC_FP32)
    stream = torch.cuda.ExternalStream(soi.stream_of(d))
    ref = torch.fft.fft(x.to(torch.complex128))
    y = x.clone()
    torch.cuda.synchronize()
    soi.compute_soi(d, y.data_ptr())
    err = (torch.linalg.norm(y.to(torch.complex128) - ref) / torch.linalg.norm(ref)).item()
    del ref, y
    torch.cuda.empty_cache()
    bufs = step_buffers(torch, x, args.steps)
    for _ in range(args.warmup):
        soi.compute_soi_async(d, bufs[0].data_ptr(), None)
    refill(torch, bufs, x)
    torch.cuda.synchronize()
    soi.enable_timing(d, True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(args.steps):
        soi.compute_soi_async(d, bufs[i % len(bufs)].data_ptr(), None)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    acc = soi.stage_times(d)
    del bufs
    soi.enable_timing(d, False)
    soi.free_soi_descriptor(d)
    gflops = 5.0 * N * math.log2(N) / (ms * 1e-3) / 1e9
    return {"value": round(gflops, 2), "unit": "GFLOP/s", "ms_per_step": round(ms, 4),
            "dtype": "f32", "rel_l2_vs_rocfft_fp64": err,
            "speedup_vs_fp64_form": round(gflops / fp64_gflops, 3),
            "stages": {"filter_sdft_ms": round(acc[0], 4), "fft_pass1_ms": round(acc[2], 4),
                       "fft_pass2_demod_ms": round(acc[3], 4)},
            "form": "complex64 data; FP32 filter contraction on MFMA (v_mfma_f32_16x16x4_f32, "
                    "exact FP32 products) and FP32 FFTs; FP64 window, twiddle and W_inv "
                    "generation; W_inv applied in FP64"}


def bench_batch(args, soi, torch, prec):
    """The batching form: args.batch_count independent transforms, N uniform over {2^16, 3*2^15,
    2^18, 5*2^16, 2^20}, P uniform over {8, 16, 32} (rejecting (2^16, 32), whose M/P < B),
    uniform[-1,1] inputs, captured into one HIP graph (each shape autotuned once) and replayed:
    one step = one graph launch transforming all of them."""
    import numpy as np
    sizes, ps = [2 ** 16, 3 * 2 ** 15, 2 ** 18, 5 * 2 ** 16, 2 ** 20], [8, 16, 32]
    rng = np.random.default_rng(2028)
    shapes = []
    while len(shapes) < args.batch_count:
        n, p = int(rng.choice(sizes)), int(rng.choice(ps))
        if (n, p) != (2 ** 16, 32):
            shapes.append((n, p))
    cdt = torch.complex64 if prec else torch.complex128
    descs, bufs = [], []
    t0 = time.perf_counter()
    for i, (n, p) in enumerate(shapes):
        descs.append(soi.make_descriptor(n, p, precision=prec))
        b = torch.empty(n, dtype=torch.complex128, device="cuda")
        soi.lib().soi_fill_uniform(b.data_ptr(), n, 7000 + i, None)
        bufs.append(b.to(cdt) if prec else b)
    torch.cuda.synchronize()
    batch = soi.Batch(descs, [b.data_ptr() for b in bufs], autotune=True)
    setup_s = time.perf_counter() - t0
    stream = torch.cuda.ExternalStream(batch.stream)
    for _ in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        batch.run()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    flop = sum(5.0 * n * math.log2(n) for n, _ in shapes)
    batch.close()
    for d in descs:
        soi.free_soi_descriptor(d)
    del bufs
    torch.cuda.empty_cache()
    return {"value": round(flop / (ms * 1e-3) / 1e9, 2), "unit": "GFLOP/s",
            "ms_per_step": round(ms, 4), "transforms": len(shapes),
            "points": int(sum(n for n, _ in shapes)), "setup_s": round(setup_s, 2),
            "dtype": "f32" if prec else "f64",
            "launch": "one HIP graph launch per step; inside it one launch per kernel per group of "
                      "same-shape transforms (blockIdx.y = transform), groups over 4 capture branches"}


def cpu_baseline(log2n, P):
    """The oracle (C restatement of the reference, OpenMP) on this host's cores, on a bounded
    sample: N = 2^log2n with the same S, mu, B as the GPU workload."""
    import numpy as np
    import oracle
    N = 1 << log2n
    p = oracle.params(N, P)
    x = oracle.uniform_input(N, 5)
    t = time.perf_counter()
    oracle.compute_soi(p, x)
    dt = time.perf_counter() - t
    return {"value": 5 * N * math.log2(N) / dt / 1e9, "unit": "GFLOP/s",
            "cores": int(oracle.lib().orc_num_threads()), "kind": "port",
            "sample": f"oracle compute_soi, N=2^{log2n} complex128, P={P}, k=1, mu=5/4, B=72 "
                      f"(same S as the GPU workload), {dt:.2f} s"}


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
    stream = torch.cuda.ExternalStream(soi.stream_of(d))
    # accuracy of the step as timed: one transform against rocFFT's FP64 FFT
    ref = torch.fft.fft(x)
    y = x.clone()
    torch.cuda.synchronize()
    soi.compute_soi(d, y.data_ptr())
    rel_l2 = (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item()
    if args.copy_calibration:
        torch.neg(x, out=y)
        torch.cuda.synchronize()
    del ref, y
    torch.cuda.empty_cache()

    bufs = step_buffers(torch, x, args.steps)
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
    flop = 5.0 * N * math.log2(N)
    gflops = flop / (ms * 1e-3) / 1e9
    # mean per-kernel durations over the timed steps (the last 64 when steps > 64)
    acc = soi.stage_times(d)
    del bufs
    soi.enable_timing(d, False)
    S, M, Mh = d.S, d.M, d.M_hat
    soi.free_soi_descriptor(d)

    mu, B = 5 / 4, 72
    n1, n2 = fft_split(Mh)
    # algorithmic bytes and flops per launch of each kernel (one launch covers all S segments)
    kernels = {
        "k_filter": (acc[0], 16.0 * N * (1 + mu),
                     (4.0 * B + 6 + 5 * math.log2(S)) * mu * N),
        "k_fft_cols": (acc[2], 32.0 * mu * N, (5.0 * math.log2(n1) + 6) * mu * N),
        "k_fft_rows": (acc[3], 16.0 * mu * N + 16.0 * N + 16.0 * M,
                       5.0 * math.log2(n2) * mu * N + 6.0 * N),
    }
    name = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes, kflops = kernels[name]
    hbm_frac = kbytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS
    fp64_frac = kflops / (kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS
    hbm = hbm_frac >= fp64_frac
    roofline = {"bound": "hbm" if hbm else "mfma",
                "achieved": round(kbytes / (kms * 1e-3) / 1e9 if hbm
                                  else kflops / (kms * 1e-3) / 1e12, 2),
                "peak": HBM_PEAK_GBS if hbm else FP64_PEAK_TFLOPS,
                "unit": "GB/s" if hbm else "TFLOP/s",
                "frac": round(max(hbm_frac, fp64_frac), 4),
                "traffic": traffic_from_profiles(f"{name} N=2^{args.log2n} P={P}"),
                "kernel": name, "launch_ms": round(kms, 4),
                "algorithmic_bytes": kbytes, "algorithmic_flops": kflops,
                "hbm_frac": round(hbm_frac, 4), "fp64_frac": round(fp64_frac, 4)}
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
                   "B": B, "form": "fp64", "fft_split": [n1, n2]},
        "rel_l2_vs_rocfft_fp64": rel_l2,
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
        out["fp32_split"] = bench_fp32(args, soi, torch, x, N, P, gflops)
    if not args.no_batch:
        del x
        torch.cuda.empty_cache()
        out["batch_form"] = {"fp64": bench_batch(args, soi, torch, 0),
                             "fp32": bench_batch(args, soi, torch, 1)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_log2n, P)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
