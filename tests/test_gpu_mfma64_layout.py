"""Lane layout of v_mfma_f64_16x16x4_f64, which the complex128 MFMA filter relies on.

Model (16 x 16 x 4, D = A B + C):
  A[i][k] in lane i + 16 k;  B[k][j] in lane j + 16 k;
  C[i][j] and D[i][j] in lane j + 16 (i % 4), register i // 4.
(The result layout differs from v_mfma_f32_16x16x4_f32's, where D[i][j] is in lane
j + 16 (i // 4), register i % 4.)
The probe kernel (soik_mfma64_layout_probe) runs 64 one-hot experiments (A = 1 at one lane,
B = lane + 1), from which the three mappings are read off independently of the model, and one
experiment with small integers in A, B and C, where every product and sum is exact in FP64."""
import ctypes

import numpy as np
import pytest

from gpu_util import d2h

pytestmark = pytest.mark.gpu


def a_pos(lane):
    return lane % 16, lane // 16          # (i, k)


def b_pos(lane):
    return lane // 16, lane % 16          # (k, j)


def d_pos(lane, r):
    return 4 * r + lane // 16, lane % 16    # (i, j)


def model_product(a, b, c):
    A = np.zeros((16, 4))
    B = np.zeros((4, 16))
    C = np.zeros((16, 16))
    for l in range(64):
        A[a_pos(l)] = a[l]
        B[b_pos(l)] = b[l]
        for r in range(4):
            C[d_pos(l, r)] = c[l, r]
    D = A @ B + C
    out = np.zeros((64, 4))
    for l in range(64):
        for r in range(4):
            out[l, r] = D[d_pos(l, r)]
    return out


@pytest.fixture(scope="module")
def probe(soi, torch_cuda):
    L = soi.lib()
    L.soik_mfma64_layout_probe.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.soik_mfma64_layout_probe.restype = ctypes.c_int
    buf = torch_cuda.zeros(65 * 256, dtype=torch_cuda.float64, device="cuda")
    torch_cuda.cuda.synchronize()
    assert L.soik_mfma64_layout_probe(buf.data_ptr(), None) == 0
    torch_cuda.cuda.synchronize()
    return buf.cpu().numpy().reshape(65, 64, 4)


def test_one_hot_experiments_match_model(probe):
    lanes = np.arange(64)
    for e in range(64):
        a = (lanes == e).astype(float)
        b = (lanes + 1).astype(float)
        want = model_product(a, b, np.zeros((64, 4)))
        got = probe[e]
        if not np.array_equal(got, want):
            nz = [(l, r, int(got[l, r])) for l in range(64) for r in range(4) if got[l, r]]
            pytest.fail(f"A lane {e}: nonzero (lane, reg, B lane + 1) = {nz}")


def test_general_experiment_matches_model(probe):
    lanes = np.arange(64)
    a = ((lanes * 7) % 11 - 5).astype(float)
    b = ((lanes * 5) % 13 - 6).astype(float)
    c = (100 * lanes[:, None] + np.arange(4)[None, :]).astype(float)
    assert np.array_equal(probe[64], model_product(a, b, c))
