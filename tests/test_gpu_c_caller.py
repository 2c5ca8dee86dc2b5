"""A reference-style C caller, run: the call sequence of test.c:199-207, 373-391, 406-420 and 460
written against include/soi.h, compiled with gcc and linked against libsoi_hip.so, on a host
buffer (staged through the device by the library) and on a device buffer (hipMalloc).  The
output is checked against the oracle at the FP64 bar.  test.c's free() calls on d.w, d.w_dup and
d.W_inv are the lines a porting caller drops (those are device buffers here); the nulling of
alpha_tilde, beta_tilde and gamma_tilde is kept."""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

CALLER_C = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>
#include "soi.h"

/* caller <host|device> <N> <P> <k> <in.bin> <out.bin>: complex128 files of N values */
int main(int argc, char **argv) {
  if (argc != 7) return 64;
  int on_device = strcmp(argv[1], "device") == 0;
  size_t N = strtoull(argv[2], 0, 10);
  soi_desc_t d;
  set_default_soi_descriptor(&d);
  d.N = N; d.P = atoi(argv[3]); d.rank = 0;
  cfft_complex_t *in_buf = (cfft_complex_t *)malloc(sizeof(cfft_complex_t) * N);
  FILE *f = fopen(argv[5], "rb");
  if (!f || fread(in_buf, sizeof(cfft_complex_t), N, f) != N) return 65;
  fclose(f);
  for (int rep = 0; rep < 2; ++rep) {          /* test.c's k loop: init, compute, null, free */
    init_soi_descriptor(&d, MPI_COMM_WORLD, (cfft_size_t)atoi(argv[4]));
    if (soi_last_status()) { fprintf(stderr, "init: %s\n", soi_last_error()); return 2; }
    cfft_complex_t *buf = in_buf, *dbuf = NULL;
    cfft_complex_t *work = (cfft_complex_t *)malloc(sizeof(cfft_complex_t) * N);
    memcpy(work, in_buf, sizeof(cfft_complex_t) * N);
    buf = work;
    if (on_device) {
      if (hipMalloc((void **)&dbuf, sizeof(cfft_complex_t) * N) != hipSuccess) return 3;
      if (hipMemcpy(dbuf, work, sizeof(cfft_complex_t) * N, hipMemcpyHostToDevice)) return 3;
      buf = dbuf;
    }
    compute_soi(&d, buf);
    if (soi_last_status()) { fprintf(stderr, "compute: %s\n", soi_last_error()); return 4; }
    if (on_device && hipMemcpy(work, dbuf, sizeof(cfft_complex_t) * N, hipMemcpyDeviceToHost))
      return 3;
    /* test.c:416-420: keep the buffers for reuse by nulling the descriptor's pointers */
    d.alpha_tilde = NULL;
    d.beta_tilde = NULL;
    d.gamma_tilde = NULL;
    int seg = d.segmentBoundaries[d.P] - d.segmentBoundaries[0];
    free_soi_descriptor(&d);
    if (soi_last_status() || seg != (int)(d.k * d.P)) return 5;
    if (rep == 1) {
      f = fopen(argv[6], "wb");
      if (!f || fwrite(work, sizeof(cfft_complex_t), N, f) != N) return 66;
      fclose(f);
    }
    if (dbuf) (void)hipFree(dbuf);
    free(work);
  }
  free(in_buf);
  return 0;
}
"""


@pytest.fixture(scope="module")
def caller(tmp_path_factory, soi):
    d = tmp_path_factory.mktemp("caller")
    src = d / "caller.c"
    src.write_text(CALLER_C)
    exe = d / "caller"
    libdir = os.path.dirname(soi.LIB_PATH)
    subprocess.check_call(["gcc", "-std=gnu11", "-D__HIP_PLATFORM_AMD__", f"-I{INC}",
                           f"-I{ROCM}/include", str(src), "-o", str(exe), f"-L{libdir}",
                           "-lsoi_hip", f"-Wl,-rpath,{libdir}", f"-L{ROCM}/lib", "-lamdhip64",
                           f"-Wl,-rpath,{ROCM}/lib", "-lm"])
    return exe


@pytest.fixture(scope="module")
def caller32(tmp_path_factory, soi):
    """The same caller built with -DPRECISION=1 (reference intrinsic.h:4-5): float complex
    buffers, and set_default_soi_descriptor selects the FP32 form."""
    d = tmp_path_factory.mktemp("caller32")
    src = d / "caller.c"
    src.write_text(CALLER_C)
    exe = d / "caller32"
    libdir = os.path.dirname(soi.LIB_PATH)
    subprocess.check_call(["gcc", "-std=gnu11", "-DPRECISION=1", "-D__HIP_PLATFORM_AMD__",
                           f"-I{INC}", f"-I{ROCM}/include", str(src), "-o", str(exe),
                           f"-L{libdir}", "-lsoi_hip", f"-Wl,-rpath,{libdir}", f"-L{ROCM}/lib",
                           "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-lm"])
    return exe


@pytest.mark.parametrize("where", ["host", "device"])
def test_precision1_c_caller_runs_the_fp32_form(caller32, orc, tmp_path, where):
    N, P, k = 2 ** 16, 8, 1
    print(f"case C caller PRECISION=1 {where} N={N} P={P}", flush=True)
    x = orc.uniform_input(N, 78).astype(np.complex64)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(fin)
    r = subprocess.run([str(caller32), where, str(N), str(P), str(k), str(fin), str(fout)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    got = np.fromfile(fout, dtype=np.complex64).astype(np.complex128)
    ref, _ = orc.compute_soi(orc.params(N, P, k), x.astype(np.complex128))
    err = rel(got, ref)
    print(f"C caller PRECISION=1 {where}: rel-L2 vs FP64 oracle {err:.3e}")
    assert err < 1e-6


@pytest.mark.parametrize("where", ["host", "device"])
@pytest.mark.parametrize("N,P,k", [(2 ** 16, 8, 1), (2 ** 18, 4, 2)])
def test_reference_style_c_caller_runs(caller, orc, tmp_path, where, N, P, k):
    print(f"case C caller {where} N={N} P={P} k={k}", flush=True)
    x = orc.uniform_input(N, 77)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    x.astype(np.complex128).tofile(fin)
    r = subprocess.run([str(caller), where, str(N), str(P), str(k), str(fin), str(fout)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    got = np.fromfile(fout, dtype=np.complex128)
    ref, _ = orc.compute_soi(orc.params(N, P, k), x)
    err = rel(got, ref)
    print(f"C caller {where} N={N} P={P} k={k}: rel-L2 vs oracle {err:.3e}")
    assert err < 1e-12
