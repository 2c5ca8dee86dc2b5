"""The command-line driver (soi_fft_amd/soi_test, the reference test.c's options and key/value
output): SNR and normalised infinity-norm error of SOI and of the rocFFT comparison leg against
the exact transforms of input kinds 0, 1 and 3 (input.c:218-281)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "soi_fft_amd", "soi_test")


def run(*args):
    r = subprocess.run([DRIVER, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    kv = {}
    for line in r.stdout.splitlines():
        if "\t" in line:
            key, val = line.split("\t", 1)
            kv[key] = float(val)
    return r.stdout, kv


def test_driver_fp64_kinds():
    out, kv = run("--P", "8", "--k_min", "1", "--k_max", "2", "--input_min", "0", "--input_max", "1",
                  "2^20")
    print(out[:2000], flush=True)
    assert out.startswith("P = 8, N = 1048576, n_mu = 5, d_mu = 4, B = 72")
    _, kv3 = run("--P", "8", "--k_min", "1", "--k_max", "2", "--input_min", "3", "--input_max", "3",
                 "2^20")
    kv.update(kv3)
    for kind in (0, 1, 3):
        for k in (1, 2):
            assert kv[f"snr_soi{kind}_{k}"] > 250, (kind, k)     # pfft.c:113 quotes 288 dB
            assert kv[f"max_err_soi{kind}_{k}"] < 1e-11
            assert kv[f"flops_soi_{k}"] > 0
        assert kv[f"snr_ref{kind}"] > 250
    assert "time_ref" in kv and "flops_ref" in kv


def test_driver_fp32_and_vlc():
    _, kv = run("--P", "8", "--k_min", "1", "--k_max", "1", "--input_min", "1", "--input_max", "1",
                "--fp32", "--no_mkl", "2^20")
    assert 120 < kv["snr_soi1_1"] < 200          # FP32 arithmetic: ~1e-7 relative error
    _, kv = run("--P", "8", "--k_min", "1", "--k_max", "1", "--input_min", "0", "--input_max", "0",
                "--vlc", "--no_mkl", "2^20")
    assert kv["snr_soi0_1"] > 250
