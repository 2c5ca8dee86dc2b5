"""The batching form: many independent transforms of uneven N and P, captured into one HIP graph
(soi_batch_create) and replayed with one launch, each shape autotuned once.

Every replay must reproduce, bit for bit, the eager compute_soi of the same descriptors (same
kernels, same variants), with eager work and synchronisation between replays; two entries are
also checked against the oracle."""
import numpy as np
import pytest

from gpu_util import rel

pytestmark = pytest.mark.gpu

SIZES = [2 ** 16, 3 * 2 ** 15, 2 ** 18, 5 * 2 ** 16, 2 ** 20]
PS = [8, 16, 32]


def draw_shapes(rng, count):
    """N uniform over SIZES, P uniform over PS, rejecting (2^16, 32) whose M/P < B."""
    out = []
    while len(out) < count:
        N, P = int(rng.choice(SIZES)), int(rng.choice(PS))
        if (N, P) == (2 ** 16, 32):
            continue
        out.append((N, P))
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_graph_replays_match_eager(soi, orc, torch_cuda, prec):
    torch = torch_cuda
    rng = np.random.default_rng(40 + prec)
    shapes = draw_shapes(rng, 12)
    print(f"case prec={prec} shapes={shapes}", flush=True)
    cdt = torch.complex64 if prec else torch.complex128
    descs, bufs, inputs = [], [], []
    batch = None
    try:
        for i, (N, P) in enumerate(shapes):
            descs.append(soi.make_descriptor(N, P, precision=prec))
            x = torch.empty(N, dtype=torch.complex128, device="cuda")
            assert soi.lib().soi_fill_uniform(x.data_ptr(), N, 100 + i, None) == 0
            x = x.to(cdt)
            inputs.append(x.clone())
            bufs.append(x)
        torch.cuda.synchronize()
        batch = soi.Batch(descs, [b.data_ptr() for b in bufs], autotune=True)
        refs = []
        for d, x0 in zip(descs, inputs):
            y = x0.clone()
            torch.cuda.synchronize()
            soi.compute_soi(d, y.data_ptr())
            refs.append(y)
        for rep in range(3):
            for b, x0 in zip(bufs, inputs):
                b.copy_(x0)
            torch.cuda.synchronize()
            batch.run()
            torch.cuda.synchronize()
            for i, (b, r) in enumerate(zip(bufs, refs)):
                assert torch.equal(b, r), f"replay {rep}, entry {i} {shapes[i]}"
            _ = float(torch.sum(torch.abs(refs[0])).item())  # eager work between replays
        # parity of two entries with the oracle (FP64 transform of the same input)
        for i in (0, len(shapes) - 1):
            N, P = shapes[i]
            x = inputs[i].to(torch.complex128).cpu().numpy()
            ref, _ = orc.compute_soi(orc.params(N, P), x)
            err = rel(bufs[i].to(torch.complex128).cpu().numpy(), ref)
            assert err < (1e-6 if prec else 1e-12), (i, shapes[i], err)
    finally:
        if batch is not None:
            batch.close()
        for d in descs:
            soi.free_soi_descriptor(d)


def test_autotune_keeps_results(soi, orc, torch_cuda):
    """Every FFT variant of a plan computes the same transform; autotune returns one of them."""
    torch = torch_cuda
    N, P = 2 ** 24, 4  # M_hat = 2560 x 2048: the most alternatives
    d = soi.make_descriptor(N, P)
    try:
        nv = soi.fft_variants(d)
        assert nv >= 2
        x = torch.empty(N, dtype=torch.complex128, device="cuda")
        assert soi.lib().soi_fill_uniform(x.data_ptr(), N, 3, None) == 0
        outs = []
        for v in range(nv):
            soi.set_fft_variant(d, v)
            y = x.clone()
            torch.cuda.synchronize()
            soi.compute_soi(d, y.data_ptr())
            outs.append(y)
        for v in range(1, nv):
            e = (torch.linalg.norm(outs[v] - outs[0]) / torch.linalg.norm(outs[0])).item()
            assert e < 1e-15, (v, e)
        scratch = x.clone()
        best = soi.autotune(d, scratch.data_ptr(), 2)
        assert 0 <= best < nv
    finally:
        soi.free_soi_descriptor(d)


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_512_member_workload(soi, orc, torch_cuda, prec):
    """The batching workload at its stated size: the exact 512-member draw bench.py times
    (seed 2028, inputs seeded 7000 + i), captured with autotuning and replayed twice.  Every
    member must equal its eager compute_soi bit for bit after each replay, and one member of
    each distinct (N, P) is checked against the oracle at the form's bar (1e-12 FP64, 1e-6 FP32
    against the FP64 transform of the same input)."""
    torch = torch_cuda
    shapes = draw_shapes(np.random.default_rng(2028), 512)
    assert len(set(shapes)) == 14
    print(f"case 512-member batch prec={prec}", flush=True)
    cdt = torch.complex64 if prec else torch.complex128
    descs, bufs, inputs = [], [], []
    batch = None
    try:
        for i, (N, P) in enumerate(shapes):
            descs.append(soi.make_descriptor(N, P, precision=prec))
            x = torch.empty(N, dtype=torch.complex128, device="cuda")
            assert soi.lib().soi_fill_uniform(x.data_ptr(), N, 7000 + i, None) == 0
            x = x.to(cdt)
            inputs.append(x.clone())
            bufs.append(x)
        torch.cuda.synchronize()
        batch = soi.Batch(descs, [b.data_ptr() for b in bufs], autotune=True)
        for rep in range(2):
            for b, x0 in zip(bufs, inputs):
                b.copy_(x0)
            torch.cuda.synchronize()
            batch.run()
            torch.cuda.synchronize()
            for i, (d, x0) in enumerate(zip(descs, inputs)):
                y = x0.clone()
                torch.cuda.synchronize()
                soi.compute_soi(d, y.data_ptr())
                assert torch.equal(bufs[i], y), f"replay {rep}, member {i} {shapes[i]}"
        seen = set()
        for i, (N, P) in enumerate(shapes):
            if (N, P) in seen:
                continue
            seen.add((N, P))
            x = inputs[i].to(torch.complex128).cpu().numpy()
            ref, _ = orc.compute_soi(orc.params(N, P), x)
            err = rel(bufs[i].to(torch.complex128).cpu().numpy(), ref)
            assert err < (1e-6 if prec else 1e-12), (i, (N, P), err)
        assert len(seen) == 14
    finally:
        if batch is not None:
            batch.close()
        for d in descs:
            soi.free_soi_descriptor(d)


def test_batch_create_keeps_timing(soi, torch_cuda):
    """Capture runs with per-stage timing off (its events would become graph nodes); the
    descriptors' own timing setting is restored afterwards."""
    torch = torch_cuda
    N, P = 2 ** 18, 8
    d = soi.make_descriptor(N, P)
    batch = None
    try:
        x = torch.empty(N, dtype=torch.complex128, device="cuda")
        assert soi.lib().soi_fill_uniform(x.data_ptr(), N, 5, None) == 0
        torch.cuda.synchronize()
        soi.enable_timing(d, True)
        batch = soi.Batch([d], [x.data_ptr()], autotune=False)
        batch.run()
        torch.cuda.synchronize()
        soi.compute_soi(d, x.data_ptr())
        t = soi.stage_times(d)
        assert len(t) == 4 and t[0] > 0 and t[2] + t[3] > 0, t
    finally:
        if batch is not None:
            batch.close()
        soi.free_soi_descriptor(d)


def test_batch_long_rows_match_eager(soi, torch_cuda):
    """Members whose segment FFTs use the sequence-fastest row pass (M_hat = 1280 x 1024 at
    N = 2^22, P = 4 and 2560 x 2048 at N = 2^24, P = 4), two of each shape in one graph: every
    member equal to its eager compute_soi bit for bit."""
    torch = torch_cuda
    shapes = [(2 ** 22, 4), (2 ** 24, 4), (2 ** 22, 4), (2 ** 24, 4)]
    print(f"case batch long rows {shapes}", flush=True)
    descs, bufs, inputs = [], [], []
    batch = None
    try:
        for i, (N, P) in enumerate(shapes):
            descs.append(soi.make_descriptor(N, P))
            x = torch.empty(N, dtype=torch.complex128, device="cuda")
            assert soi.lib().soi_fill_uniform(x.data_ptr(), N, 900 + i, None) == 0
            inputs.append(x.clone())
            bufs.append(x)
        torch.cuda.synchronize()
        batch = soi.Batch(descs, [b.data_ptr() for b in bufs], autotune=False)
        batch.run()
        torch.cuda.synchronize()
        for i, (d, x0) in enumerate(zip(descs, inputs)):
            y = x0.clone()
            torch.cuda.synchronize()
            soi.compute_soi(d, y.data_ptr())
            torch.cuda.synchronize()
            assert torch.equal(bufs[i], y), f"member {i} {shapes[i]}"
    finally:
        if batch is not None:
            batch.close()
        for d in descs:
            soi.free_soi_descriptor(d)
