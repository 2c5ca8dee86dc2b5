"""Pin the CPU oracle (oracle/soi_oracle.c) before it is trusted as the checker.

Every oracle function is checked against something it does not share code with:
  * orc_fft            long-double brute-force DFT (small L), numpy.fft (large L)
  * orc_filter_subsample  long-double brute force of the Step-1 matrix-vector product
  * orc_compute_soi    numpy.fft.fft and the closed forms of the reference's input kinds 0/1
                       (input.c:218-231, 256-281) within the SOI truncation error
  * windows            an independent numpy evaluation of the closed forms (pfft.c:42-82)
  * VLC codec          the reference's known-answer test (compress.c:546-597), an independent
                       string-based bit packer, rounding-mode probes and round-trip bounds
  * load balance       hand-worked cases of pfft.c:390-440
"""
import numpy as np
import pytest
from scipy.special import erfc

from closed_forms import kind0_input, kind0_output, tones_dft, tones_input


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


# ------------------------------------------------------------------------------------------
# FFT
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("L", [1, 2, 8, 15, 60, 128, 160, 240, 320, 1000, 1536, 2560, 231])
def test_fft_vs_bruteforce(orc, L):
    x = orc.uniform_input(L, 11)
    assert rel(orc.fft(x), orc.dft_brute(x)) < 1e-15 * max(1, np.log2(L))


@pytest.mark.parametrize("L", [20480, 163840])
def test_fft_vs_numpy(orc, L):
    x = orc.uniform_input(L, 12)
    assert rel(orc.fft(x), np.fft.fft(x)) < 2e-15


def test_fft_plants_detected(orc):
    # a sign or index mistake must be visible: conj transform and a shifted input both fail
    x = orc.uniform_input(64, 3)
    ref = orc.dft_brute(x)
    assert rel(np.conj(orc.fft(np.conj(x))), ref) > 1e-3
    assert rel(orc.fft(np.roll(x, 1)), ref) > 1e-3


# ------------------------------------------------------------------------------------------
# Windows (pfft.c:42-82)
# ------------------------------------------------------------------------------------------
def _w_closed(p):
    S, M, Mh = orc_dims(p)
    th = np.arange(p.n_mu)[:, None, None]
    b = np.arange(p.B)[None, :, None]
    i = np.arange(S)[None, None, :]
    kappa = p.B - p.d_mu
    # exact rational t = theta d/n - b - i/S + kappa/2 over the common denominator 2 n S
    num = 2 * th * p.d_mu * S - 2 * b * p.n_mu * S - 2 * i * p.n_mu + kappa * p.n_mu * S
    t = num / (2.0 * p.n_mu * S)
    with np.errstate(invalid="ignore", divide="ignore"):
        y = np.where(num == 0, 1.0,
                     np.sin(np.pi * p.tau * t) / (np.pi * p.tau * t) * np.exp(-np.pi ** 2 * t * t / p.sigma))
    half = np.mod(num, 4 * p.n_mu * S) / (2.0 * p.n_mu * S)  # t mod 2
    return np.exp(1j * np.pi * half) * y * p.d_mu / p.n_mu


def orc_dims(p):
    S = p.k * p.P
    M = p.N // S
    return S, M, p.n_mu * M // p.d_mu


@pytest.mark.parametrize("shape", [(2 ** 13, 4, 1, 5, 4), (2 ** 16, 8, 2, 5, 4), (7 * 2 ** 12, 4, 1, 8, 7)])
def test_window_table(orc, shape):
    N, P, k, n, d = shape
    p = orc.params(N, P, k, n_mu=n, d_mu=d)
    w = orc.w_table(p)
    ref = _w_closed(p)
    err = np.abs(w - ref).max() / np.abs(ref).max()
    assert err < 2e-14
    # structure: |w| is the real window y(t)/mu, symmetric about t = 0
    assert abs(np.abs(w).max() - d / n) < 1e-15


def test_w_inv_table(orc):
    p = orc.params(2 ** 16, 8)
    S, M, Mh = orc_dims(p)
    W = orc.w_inv_table(p)
    m = np.arange(M)
    t = m / M - 0.5
    y = (erfc(np.sqrt(p.sigma) * (t - p.tau / 2)) - erfc(np.sqrt(p.sigma) * (t + p.tau / 2))) / (2 * p.tau)
    kappa = p.B - p.d_mu
    ref = np.exp(-1j * np.pi * np.mod(kappa * m, 2 * M) / M) / y
    assert rel(W, ref) < 1e-15
    # |W_inv| range quoted for the defaults: about [0.906, 9.06], max at m = 0
    assert 0.9 < np.abs(W).min() < 0.91 and 9.0 < np.abs(W).max() < 9.1
    assert np.argmax(np.abs(W)) == 0


# ------------------------------------------------------------------------------------------
# Step 1 and the whole transform
# ------------------------------------------------------------------------------------------
def test_filter_vs_bruteforce(orc):
    p = orc.params(2 ** 13, 4)
    x = orc.uniform_input(p.N, 3)
    fo = orc.filter_subsample(p, x)
    fb = orc.filter_brute(p, x)
    assert rel(fo, fb) < 3e-14


def test_filter_vs_bruteforce_k2(orc):
    p = orc.params(2 ** 13, 2, k=2)
    x = orc.uniform_input(p.N, 4)
    assert rel(orc.filter_subsample(p, x), orc.filter_brute(p, x)) < 3e-14


@pytest.mark.parametrize("shape", [
    (2 ** 13, 4, 1), (2 ** 16, 8, 1), (2 ** 16, 4, 2), (2 ** 20, 8, 1), (3 * 2 ** 14, 4, 1),
    (5 * 2 ** 14, 8, 1)])
def test_soi_vs_numpy_fft(orc, shape):
    N, P, k = shape
    p = orc.params(N, P, k)
    x = orc.uniform_input(N, 5)
    out, sb = orc.compute_soi(p, x)
    assert rel(out, np.fft.fft(x)) < 3e-14
    assert list(sb) == [q * k for q in range(P + 1)]


@pytest.mark.parametrize("row", [  # pfft.c:111-137 table rows (mu, tau, sigma, B, SNR dB)
    (5, 4, 872 / 1024, 313.1715, 76, 289.2718),
    (5, 4, 818 / 1024, 267.4513, 64, 284.7614),
    (5, 4, 512 / 1024, 121.0936, 44, 241.8972),
])
def test_soi_accuracy_table(orc, row):
    """Windows evaluated exactly reach the published SNR of the wide windows (B >= 64; the B = 44
    row measures about 226 dB on uniform input against 241.9 published); forming t in double as
    pfft.c:52 does adds an error floor near 1e-14 (about 270-279 dB)."""
    n, d, tau, sigma, B, snr_db = row
    p = orc.params(2 ** 16, 8, n_mu=n, d_mu=d, B=B, tau=tau, sigma=sigma)
    x = orc.uniform_input(p.N, 9)
    ref = np.fft.fft(x)

    def snr_of(out):
        return 10 * np.log10(np.sum(np.abs(ref) ** 2) / np.sum(np.abs(out - ref) ** 2))

    try:
        orc.set_window_mode(1)
        exact = snr_of(orc.compute_soi(p, x)[0])
    finally:
        orc.set_window_mode(0)
    literal = snr_of(orc.compute_soi(p, x)[0])
    assert exact > snr_db - (3 if B >= 64 else 20)
    assert literal > min(snr_db, 275) - 20
    assert exact >= literal - 1


def test_soi_kind0_closed_form(orc):
    N = 2 ** 16
    p = orc.params(N, 8)
    out, _ = orc.compute_soi(p, kind0_input(N))
    ref = kind0_output(N)
    snr = 10 * np.log10(np.sum(np.abs(ref) ** 2) / np.sum(np.abs(out - ref) ** 2))
    assert snr > 270


def test_soi_tone_closed_form(orc):
    # kind 1 (e^{i n theta}, input.c:218-224) with theta on a representable bin grid
    N = 2 ** 16
    p = orc.params(N, 8)
    f = [N / (2 * np.pi), 1234.625]
    ph = [0.0, 0.3]
    a = [1.0, 0.5]
    out, _ = orc.compute_soi(p, tones_input(N, f, ph, a))
    ref = tones_dft(N, f, ph, a, np.arange(N))
    assert rel(out, ref) < 3e-14


MU87_TAU, MU87_SIGMA = 0.7, 360.0   # searched for mu = 8/7, B = 72: about 228 dB


def test_soi_mu_8_7(orc):
    # pfft.c:111-137 tabulates no window for mu = 8/7; this one was searched with the oracle
    N = 7 * 2 ** 13
    p = orc.params(N, 4, n_mu=8, d_mu=7, B=72, tau=MU87_TAU, sigma=MU87_SIGMA)
    x = orc.uniform_input(N, 6)
    out, _ = orc.compute_soi(p, x)
    assert rel(out, np.fft.fft(x)) < 2e-11


# ------------------------------------------------------------------------------------------
# VLC codec (compress.c)
# ------------------------------------------------------------------------------------------
KAT_SEG = [["0X1.0123456789ABCp8", "0X1.DEF0123456789p8", "-0X1.0123456789ABCp8", "-0X1.DEF0123456789p8"],
           ["0X1.0123456789ABCp0", "0X1.DEF0123456789p-1", "-0X1.0123456789ABCp0", "-0X1.DEF0123456789p0"]]
KAT_EXP = [["0X1.0123456789ABCp8", "0X1.DEF0123456789p8", "-0X1.0123456789ABCp8", "-0X1.DEF0123456789p8"],
           ["0X1.0123456789Ap0", "0X1.DEF01234566p-1", "-0X1.0123456789Bp0", "-0X1.DEF01234568p0"]]


def test_vlc_known_answer(orc):
    """compress.c:546-597: two 4-element segments; the expected values follow floor rounding."""
    seg = [np.array([float.fromhex(h) for h in s]) for s in KAT_SEG]
    exp = [np.array([float.fromhex(h) for h in s]) for s in KAT_EXP]
    em = [orc.max_exponent(s) for s in seg]
    assert em == [8, 0]
    g = max(em)
    for s in range(2):
        words = orc.compress(seg[s], g, em[s], floor_mode=True)
        out = orc.decompress(words, 4, g, em[s])
        assert np.array_equal(out, exp[s])
    # the live code rounds to nearest: the two positive values of segment 1 round up
    words = orc.compress(seg[1], g, em[1])
    out = orc.decompress(words, 4, g, em[1])
    assert out[0] == float.fromhex("0X1.0123456789Bp0") and out[2] == exp[1][2]


def test_max_exponent_cases(orc):
    assert orc.max_exponent(np.array([1.0])) == 0
    assert orc.max_exponent(np.array([0.5, -3.0])) == 1
    assert orc.max_exponent(np.zeros(5)) == -1023
    assert orc.max_exponent(np.array([5e-324, 0.0])) == -1023        # denormal: field 0
    assert orc.max_exponent(np.array([np.inf])) == 1024
    x = np.full(19, 0.25)
    x[17] = -70.0                                                     # in the scalar tail
    assert orc.max_exponent(x) == 6


def _pack_reference_layout(i1, i2, i1len, i2len, nvec):
    """Independent packer: fields as bit strings, 16 lanes of MSB-first 32-bit words, then a
    sequential tail stream (layout of compress.c:170-203, 324-369)."""
    def bits(v, n):
        return format(v & ((1 << n) - 1), "0%db" % n) if n > 0 else ""
    words = []
    lanes = []
    for lane in range(16):
        s = "".join(bits(i1[f * 16 + lane], i1len) + bits(i2[f * 16 + lane], i2len)
                    for f in range(nvec))
        lanes.append(s)
    nrows = (len(lanes[0]) + 31) // 32
    for r in range(nrows):
        for lane in range(16):
            chunk = lanes[lane][32 * r: 32 * r + 32].ljust(32, "0")
            words.append(int(chunk, 2))
    tail = "".join(bits(i1[j], i1len) + bits(i2[j], i2len) for j in range(nvec * 16, len(i1)))
    for r in range((len(tail) + 31) // 32):
        words.append(int(tail[32 * r: 32 * r + 32].ljust(32, "0"), 2))
    return np.array([w - (1 << 32) if w >= 1 << 31 else w for w in words], dtype=np.int64)


@pytest.mark.parametrize("length,delta", [(16, 0), (35, 0), (64, 5), (100, 23), (48, 30), (21, 40)])
def test_vlc_bit_layout(orc, length, delta):
    rng = np.random.default_rng(length + delta)
    e_max = 3
    e_i = e_max - delta
    x = rng.uniform(-1, 1, length) * 2.0 ** (e_i + 1) * 0.999
    words = orc.compress(x, e_max, e_i)
    i1len = 24 - delta
    i2len = min(32, 55 - delta)
    e1, e2, e3 = 2.0 ** (21 - e_max), 2.0 ** (52 - e_max), 2.0 ** 31
    nvec = length // 16
    i1, i2 = [], []
    for j, v in enumerate(x):
        if j < nvec * 16:   # vector body: round half even, fused x*e2 - d1*e3
            d1 = np.rint(v * e1)
        else:               # scalar tail: round half away from zero
            d1 = np.sign(v * e1) * np.floor(abs(v * e1) + 0.5)
        r2 = v * e2 - d1 * e3   # exact (power-of-two scalings of doubles within range)
        d2 = np.rint(r2) if j < nvec * 16 else np.sign(r2) * np.floor(abs(r2) + 0.5)
        i1.append(int(d1))
        i2.append(int(d2))
    ref = _pack_reference_layout(i1, i2, i1len, i2len, nvec)
    assert np.array_equal(words.astype(np.int64), ref)


def test_vlc_rounding_modes(orc):
    # x*e1 exactly at a half: vector body rounds to even, tail rounds away from zero
    e_max = 0
    e1 = 2.0 ** 21
    x = np.zeros(17)
    x[:] = 2.5 / e1          # d1 = 2.5 everywhere
    words = orc.compress(x, e_max, e_max)
    vec_first = int(words[0]) & 0xFFFFFFFF
    assert (vec_first >> 8) == 2                          # 24-bit i1 field of lane 0: rint -> 2
    tail_word = int(words[-2]) & 0xFFFFFFFF                    # tail: 56 bits over two words
    assert (tail_word >> 8) == 3                          # round(2.5) = 3


@pytest.mark.parametrize("delta", [0, 3, 23, 24, 30, 54])
def test_vlc_round_trip_bound(orc, delta):
    rng = np.random.default_rng(delta)
    e_max = 5
    e_i = e_max - delta
    x = rng.uniform(-1, 1, 1000) * 2.0 ** e_i
    words = orc.compress(x, e_max, e_i)
    y = orc.decompress(words, len(x), e_max, e_i)
    # fixed point with lsb 2^(e_max-52): error at most half an lsb
    assert np.max(np.abs(x - y)) <= 2.0 ** (e_max - 53) * 1.0000001


def test_vlc_nothing_sent(orc):
    x = np.ones(32) * 2.0 ** -60
    assert len(orc.compress(x, 0, -60)) == 0
    assert np.all(orc.decompress(np.zeros(0, dtype=np.int32), 32, 0, -60) == 0)


def test_vlc_plan_balanced(orc):
    sc, sb = orc.vlc_plan(4, 2, 1024, 4096, 1.0, [3] * 8)
    assert list(sb) == [0, 2, 4, 6, 8]
    # 56 bits per input, 2048 doubles: 128 per lane * 56 / 32 = 224 words per lane
    assert all(c == (224 * 16 + 1) // 2 for c in sc)


def test_vlc_plan_skewed(orc):
    # segments far below the global maximum send little; ranks then take more of them
    g = [10, 10, 10, 10, -40, -40, -40, -40]
    sc, sb = orc.vlc_plan(2, 4, 1 << 14, 1 << 15, 50.0, g)
    assert sb[0] == 0 and sb[2] == 8
    assert 0 < sb[1] < 4            # the expensive segments are split over fewer rank-0 slots
    assert sc[4] < sc[0]


def test_vlc_soi_changes_little(orc):
    p = orc.params(2 ** 14, 4)
    x = orc.uniform_input(p.N, 8)
    a, _ = orc.compute_soi(p, x)
    b, sb = orc.compute_soi(p, x, use_vlc=True)
    assert rel(b, a) < 1e-14
    assert sb[0] == 0 and sb[-1] == p.P
