"""Closed-form inputs and exact DFTs used as independent checks.

Kinds 0 and 1 restate the reference's known-answer inputs (input.c:218-231) and their exact
transforms (input.c:256-281).  Tones generalise kind 1 to fractional frequencies.
"""
import numpy as np


def kind1_input(N, theta=1.0):
    """e^{i n theta}, evaluated in long double like the reference's cosl/sinl (input.c:218-224)."""
    a = np.arange(N, dtype=np.longdouble) * np.longdouble(theta)
    return (np.cos(a) + 1j * np.sin(a)).astype(np.complex128)


def kind1_output(N, theta=1.0, idx=None):
    """sum_n e^{i n theta} e^{-2 pi i n k/N} = (1 - e^{i N theta}) / (1 - e^{i (theta - 2 pi k/N)})."""
    k = np.arange(N, dtype=np.float64) if idx is None else np.asarray(idx, dtype=np.float64)
    num = 1 - np.cos(theta * N) - 1j * np.sin(theta * N)
    phi = theta - 2 * np.pi * k / N
    s = np.sin(phi / 2)
    den = 2 * s * s - 1j * np.sin(phi)
    return num / den


def kind0_input(N):
    """Rectangle of width W = N/16 + 1 centred on 0 (circularly), height 1/W."""
    W = N // 16 + 1
    n = np.arange(N)
    return np.where((2 * n < W) | (2 * (N - n) < W), 1.0 / W, 0.0).astype(np.complex128)


def kind0_output(N):
    """Dirichlet kernel sin(pi W k/N) / (W sin(pi k/N)) with exact range reduction of W k/N."""
    W = N // 16 + 1
    k = np.arange(N, dtype=np.int64)
    L = (k * W + N // 2) // N                      # round(k W / N)
    f = (k * W - L * N).astype(np.float64) / N     # in [-1/2, 1/2)
    numer = np.where(L & 1, -np.sin(np.pi * f), np.sin(np.pi * f))
    kk = np.where(2 * k <= N, k, N - k).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        out = numer / (W * np.sin(np.pi * kk / N))
    out[0] = 1.0
    return out.astype(np.complex128)


def tones_input(N, freqs, phases, amps):
    """sum_j a_j e^{2 pi i (f_j n/N + phi_j)} with the phase reduced exactly in turns."""
    n = np.arange(N, dtype=np.int64)
    x = np.zeros(N, dtype=np.complex128)
    for f, ph, a in zip(freqs, phases, amps):
        fi = np.floor(f)
        ff = f - fi
        turns = ((int(fi) % N) * n % N) / N + ff * (n / N) + ph
        turns -= np.floor(turns)
        x += a * np.exp(2j * np.pi * turns)
    return x


def tones_dft(N, freqs, phases, amps, bins):
    """Exact DFT at the given bins of sum_j a_j e^{2 pi i (f_j n/N + phi_j)}.
    With d = f - k reduced exactly to dd = d - rN (r = round(d/N)) and dn = dd - round(dd):
      sum_n e^{2 pi i d n/N} = e^{i pi (dn - dd/N)} sin(pi dn) / sin(pi dd/N)
    (the (-1)^round factors of numerator and phase cancel), and N when dd = 0."""
    k = np.asarray(bins, dtype=np.int64)
    out = np.zeros(k.shape, dtype=np.complex128)
    for f, ph, a in zip(freqs, phases, amps):
        fi = int(np.floor(f))
        F = f - fi                                   # exact, in [0, 1)
        I = fi - k                                   # exact integer part of d = f - k
        r = np.round((I + F) / N).astype(np.int64)
        dd_int = I - r * N
        dn = F - np.round(F)                         # d - round(d), exact
        e = (dd_int + F) / N
        dd = dd_int + F
        s_den = np.sin(np.pi * e)
        hit = dd == 0
        with np.errstate(divide="ignore", invalid="ignore"):
            g = np.exp(1j * np.pi * (dn - e)) * np.sin(np.pi * dn) / s_den
        g = np.where(hit, N, g)
        out += a * np.exp(2j * np.pi * ph) * g
    return out
