"""The C ABI: libsoi_hip.so loads, exports every function include/*.h declares, the descriptor
layout seen from C matches the ctypes mirror, a reference-style C caller compiles and links, and
the checked error paths (which the reference turns into exit(), parallel_filter_subsampling.cpp:
89-94, 537-548; pfft.c:157-183; test.c:130-135) report codes instead.  No GPU needed."""
import ctypes
import os
import re
import subprocess

import pytest

import soi_fft_amd as soi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for h in sorted(os.listdir(INC)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#.*", "", text)
        text = text.replace('extern "C" {', "")
        while re.search(r"\{[^{}]*\}", text):
            text = re.sub(r"\{[^{}]*\}", "", text)
        text = re.sub(r"typedef[^;]*;", "", text)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{)]*\)\s*;", text):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_all_declared_symbols():
    L = soi.lib()
    names = declared_functions()
    assert "compute_soi" in names and "compress" in names and len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "soi.h"
#define F(f) printf(#f " %zu\n", offsetof(soi_desc_t, f));
int main(void) {
  printf("size %zu\n", sizeof(soi_desc_t));
  F(comm) F(P) F(rank) F(k) F(N) F(W_inv) F(w) F(w_dup) F(gamma_tilde) F(alpha_tilde)
  F(beta_tilde) F(alpha_ghost) F(delta) F(epsilon) F(n_mu) F(d_mu) F(B) F(tau) F(sigma)
  F(plan) F(use_vlc) F(segmentBoundaries) F(comm_to_comp_cost_ratio) F(precision)
  F(window_mode)
  return 0;
}
"""


def test_descriptor_layout_matches_ctypes(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=gnu11", f"-I{INC}", str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    got = dict(l.split() for l in out if l.strip())
    assert int(got["size"]) == ctypes.sizeof(soi.SoiDesc)
    for name, off in got.items():
        if name == "size":
            continue
        assert getattr(soi.SoiDesc, name).offset == int(off), name


# The call sequence of the reference driver (test.c:199-207, 373-391, 460), written against the
# installed header and linked against the library: a C caller switches by relinking.
CALLER_C = r"""
#include <stdio.h>
#include <stdlib.h>
#include "soi.h"
#include "compress.h"
int main(int argc, char **argv) {
  soi_desc_t d;
  set_default_soi_descriptor(&d);
  d.N = 1 << 20; d.P = 8; d.rank = 0;
  init_soi_descriptor(&d, MPI_COMM_WORLD, 1);
  if (soi_last_status()) { printf("init: %s\n", soi_last_error()); return 2; }
  cfft_complex_t *buf = (cfft_complex_t *)malloc(sizeof(cfft_complex_t) * d.N);
  for (size_t i = 0; i < d.N; ++i) buf[i] = (double)(i % 7) - 3.0;
  compute_soi(&d, buf);
  int seg0 = d.segmentBoundaries[0];
  free_soi_descriptor(&d);
  free(buf);
  return soi_last_status() || seg0;
}
"""


def test_reference_style_c_caller_links(tmp_path):
    src = tmp_path / "caller.c"
    src.write_text(CALLER_C)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(soi.LIB_PATH)
    subprocess.check_call(["gcc", "-std=gnu11", f"-I{INC}", str(src), "-o", str(exe),
                           f"-L{libdir}", "-lsoi_hip", f"-Wl,-rpath,{libdir}", "-lm"])
    assert exe.exists()


PREC1_C = r"""
#include <stdio.h>
#include "soi.h"
int main(void) {
  soi_desc_t d;
  set_default_soi_descriptor(&d);
  printf("%d %zu %d\n", PRECISION, sizeof(cfft_complex_t), d.precision);
  return 0;
}
"""


@pytest.mark.parametrize("prec,size,form", [(None, 16, 0), (2, 16, 0), (1, 8, 1)])
def test_precision_macro_maps_onto_the_forms(tmp_path, prec, size, form):
    """reference intrinsic.h:4-5: a caller built with -DPRECISION=1 gets float complex buffers
    and set_default_soi_descriptor selects the FP32 form; PRECISION 2 (or none) stays the FP64
    drop-in.  The caller runs here: set_default_soi_descriptor makes no GPU call."""
    src = tmp_path / "prec.c"
    src.write_text(PREC1_C)
    exe = tmp_path / "prec"
    libdir = os.path.dirname(soi.LIB_PATH)
    flags = [] if prec is None else [f"-DPRECISION={prec}"]
    subprocess.check_call(["gcc", "-std=gnu11", *flags, f"-I{INC}", str(src), "-o", str(exe),
                           f"-L{libdir}", "-lsoi_hip", f"-Wl,-rpath,{libdir}", "-lm"])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    p, sz, fm = (int(t) for t in out.stdout.split())
    assert (p, sz, fm) == (prec or 2, size, form)


def _desc(**kw):
    d = soi.SoiDesc()
    soi.set_default_soi_descriptor(d)
    d.N, d.P = kw.pop("N", 1 << 16), kw.pop("P", 8)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def test_defaults_match_reference():
    d = _desc()
    # pfft.c:87-94
    assert (d.n_mu, d.d_mu, d.B, d.use_vlc) == (5, 4, 72, 0)
    assert d.tau == 928 / 1024.0 and d.sigma == 373.7314 and d.comm_to_comp_cost_ratio == 1
    assert not d.alpha_tilde and not d.gamma_tilde and not d.beta_tilde


@pytest.mark.parametrize("kw,k,status", [
    (dict(n_mu=6, d_mu=5), 1, 3),            # parallel_filter_subsampling.cpp:537-548
    (dict(N=1 << 12, P=8), 1, 2),            # M/P < B: parallel_filter_subsampling.cpp:89-94
    (dict(N=1 << 16, P=6), 1, 1),            # S % 4 (pfft.c:192)
    (dict(N=1000, P=8), 1, 1),               # S does not divide N
    (dict(rank=1), 1, 1),                    # one GPU runs every rank
    (dict(), 2000, 1),                       # k <= 1024 (pfft.c:274)
    (dict(B=120), 1, 1),                     # taps beyond the supported 80
    (dict(precision=1, use_vlc=1), 1, 1),    # VLC words encode FP64 values (compress.c)
    (dict(precision=7), 1, 1),               # unknown precision
])
def test_init_rejects(kw, k, status):
    d = _desc(**kw)
    soi.lib().init_soi_descriptor(ctypes.byref(d), 0, k)
    assert soi.lib().soi_last_status() == status, soi.lib().soi_last_error()
    assert not d.plan


def test_state_errors():
    d = _desc()
    with pytest.raises(soi.SoiError) as e:
        soi.compute_soi(d, 0x1000)
    assert e.value.status == 6
    with pytest.raises(soi.SoiError):
        soi.free_soi_descriptor(d)


def test_batch_rejects_bad_arguments():
    """soi_batch_create validates before any device work: an uninitialised descriptor is a state
    error, an empty batch an argument error."""
    import ctypes
    L = soi.lib()
    d = _desc()
    P = ctypes.POINTER(soi.SoiDesc)
    arr = (P * 1)(ctypes.pointer(d))
    bufs = (ctypes.c_void_p * 1)(None)
    assert not L.soi_batch_create(arr, bufs, 1, 0)
    assert L.soi_last_status() == 6
    assert not L.soi_batch_create(arr, bufs, 0, 0)
    assert L.soi_last_status() == 1
    assert L.soi_fft_variants(ctypes.byref(d)) == 0 and L.soi_last_status() == 6


def test_driver_rejects_bad_arguments():
    """The command-line driver checks its arguments before touching the GPU (test.c:130)."""
    import subprocess
    drv = os.path.join(os.path.dirname(soi.LIB_PATH), "soi_test")
    assert os.path.exists(drv)
    r = subprocess.run([drv, "--P", "8", "--input_min", "2", "--input_max", "2", "2^20"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "input kind 2" in r.stderr
    r = subprocess.run([drv, "--P", "8", "1000"], capture_output=True, text=True)
    assert r.returncode == 1 and "must divide N" in r.stderr
    r = subprocess.run([drv], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
