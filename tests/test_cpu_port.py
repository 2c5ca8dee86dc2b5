"""The CPU performance port (oracle/soi_cpu_port.c), which bench.py times as the CPU baseline.

It is not the checker; it is checked here against the oracle (orc_compute_soi, the restatement
of pfft.c:276-721) and against numpy's FFT, on the shapes that exercise each of its code paths:
the AVX2 conv (S a multiple of 8) and the scalar conv (S = 4), radices 2/4/5/3 in the segment
FFT, k > 1, and mu = 8/7.  Both compute in FP64 from the same window tables, so they agree to
rounding; the gain of the window's demodulation (max |W_inv|) scales that rounding."""
import numpy as np
import pytest

import oracle

SHAPES = [
    (2 ** 16, 8, 1, {}),
    (5 * 2 ** 14, 8, 1, {}),
    (3 * 2 ** 14, 4, 1, {}),
    (2 ** 18, 32, 1, {}),
    (2 ** 16, 4, 2, {}),
    (7 * 2 ** 15, 8, 1, dict(n_mu=8, d_mu=7, tau=0.7, sigma=360.0)),
]


@pytest.mark.parametrize("N,P,k,kw", SHAPES)
def test_cpu_port_matches_oracle(N, P, k, kw):
    p = oracle.params(N, P, k, **kw)
    x = oracle.uniform_input(N, 3)
    ref, _ = oracle.compute_soi(p, x)
    port = oracle.CpuPort(p)
    try:
        y = port.compute(x.copy())
    finally:
        port.close()
    gain = float(np.abs(oracle.w_inv_table(p)).max())
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    assert err < max(1e-14, 30 * 2.0 ** -52 * gain), err
    # and it is the DFT, within the window's truncation error
    ex = np.fft.fft(x)
    bound = 3e-14 if kw.get("n_mu", 5) == 5 else 1e-10
    assert np.linalg.norm(y - ex) / np.linalg.norm(ex) < bound


def test_cpu_port_rejects_unsupported_length():
    # M_hat with a factor 7 has no radix in the port's FFT
    p = oracle.params(7 * 2 ** 14, 4, 1)
    with pytest.raises(ValueError):
        oracle.CpuPort(p)
