"""The quad coder with the outlier split (entries beyond the band coded as 0 and listed
per row): forward bit-exact against the integer product of W (one matrix and batched, 1
and 8 vectors, the rows' records spread over the lanes and past the prefetched slots),
backward against dense FP64 (one matrix and batched, bitwise equal and reproducible)."""
import sys

import numpy as np
import pytest

from matvans_amd.batch import Batch, BatchBackward
from matvans_amd.mvx import CompressedMatrixV1

from .helpers import gaussian_int8, laplace_int8

pytestmark = pytest.mark.gpu


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def exact(m, v):
    return (np.asarray(v, np.float64) @ np.asarray(m, np.float64).T).astype(np.int32)


def vecs(cols, n, seed):
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(rng.normal(0, 30, (n, cols))), -127, 127).astype(np.int8)


CASES = [  # (rows, cols, kind, band): band None = the library's choice
    (4096, 14336, "laplace", None), (14336, 4096, "laplace", None),
    (77, 2048, "gauss", None), (300, 2048, "gauss", 2), (5, 4096, "laplace", 0)]


def matrix(rows, cols, kind, seed):
    f = laplace_int8 if kind == "laplace" else gaussian_int8
    return f(rows, cols, 4.0 if kind == "laplace" else 6.0, seed=seed)[0]


@pytest.mark.parametrize("coder", [1, 2])
@pytest.mark.parametrize("S", [32, 64, 128])
@pytest.mark.parametrize("rows,cols,kind,band", CASES)
def test_forward_with_outliers(cuda, S, rows, cols, kind, band, coder):
    import torch
    if cols % (16 * S):
        pytest.skip("shape does not fit the stream count")
    note(f"outliers forward coder {coder} S={S} {rows}x{cols} {kind} band={band}")
    m = matrix(rows, cols, kind, rows + cols)
    c = CompressedMatrixV1.encode(m, S, 8, 0, coder=coder, n_threads=16,
                                  outliers="auto" if band is None else band).to_device()
    assert c.n_outliers > 0
    per_row = np.diff(c.out_start).max()
    for n_vec in (1, 8):
        x = vecs(cols, n_vec, 5 * n_vec)
        got = c.multiply(torch.from_numpy(x).cuda())
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), exact(m, x)), (n_vec, per_row)


def test_batch_with_outliers(cuda):
    """Quad items with and without the split and a .bin v0 item in one launch."""
    import torch
    import matvans_amd as mv
    note("outliers batch")
    specs = [(4096, 4096, 32, "auto"), (1024, 4096, 64, "auto"), (512, 2048, 32, None),
             (300, 2048, 128, 2), (256, 1024, 0, None)]
    entries, checks = [], []
    for i, (rows, cols, S, out) in enumerate(specs):
        m = matrix(rows, cols, "laplace", 40 + i)
        x = vecs(cols, 1 + i % 4, 60 + i)
        if S == 0:
            c = mv.CompressedMatrix.encode(m).to_device()
        else:
            c = CompressedMatrixV1.encode(m, S, 8, 0, coder=1, outliers=out).to_device()
        entries.append((c, torch.from_numpy(x).cuda()))
        checks.append(exact(m, x))
    b = Batch(entries)
    for cfg in [None, (2, 1, 4, 0), (1, 2, 8, 1)]:
        if cfg is not None:
            b.configure(*cfg)
        for o in b.outputs:
            o.fill_(-1)
        b.launch()
        torch.cuda.synchronize()
        for i, (o, w) in enumerate(zip(b.outputs, checks)):
            assert np.array_equal(o.cpu().numpy(), w), (i, cfg)


def dense(w, s_w, g, alpha):
    return alpha[:, None].astype(np.float64) * ((g.astype(np.float64) * s_w[None, :]) @
                                                w.astype(np.float64))


@pytest.mark.parametrize("coder", [1, 2])
@pytest.mark.parametrize("S", [32, 64])
@pytest.mark.parametrize("rows,cols,kind,band,n_vec", [
    (4096, 14336, "laplace", None, 1), (4096, 14336, "laplace", None, 8),
    (14336, 4096, "laplace", None, 2), (300, 2048, "gauss", 2, 3)])
def test_backward_with_outliers(cuda, S, rows, cols, kind, band, n_vec, coder):
    import torch
    note(f"outliers backward coder {coder} S={S} {rows}x{cols} band={band} n_vec={n_vec}")
    m = matrix(rows, cols, kind, 3 * rows + cols)
    c = CompressedMatrixV1.encode(m, S, 8, 0, coder=coder, n_threads=16,
                                  outliers="auto" if band is None else band).to_device()
    rng = np.random.default_rng(9)
    g = rng.standard_normal((n_vec, rows)).astype(np.float32)
    s_w = (rng.random(rows) * 0.02 + 0.001).astype(np.float32)
    alpha = (rng.random(n_vec) + 0.5).astype(np.float32)
    dg, ds, da = (torch.from_numpy(a).cuda() for a in (g, s_w, alpha))
    dx = c.backward(dg, ds, da)
    dx2 = c.backward(dg, ds, da)
    torch.cuda.synchronize()
    got = dx.cpu().numpy()
    assert got.tobytes() == dx2.cpu().numpy().tobytes()       # bitwise reproducible
    ref = dense(m, s_w, g, alpha)
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    bb = BatchBackward([(c, dg, ds, da)])
    bb.launch()
    torch.cuda.synchronize()
    assert bb.outputs[0].cpu().numpy().tobytes() == got.tobytes()


def test_autograd_with_outliers(cuda):
    import torch
    import matvans_amd as mv
    note("outliers autograd")
    w = matrix(512, 2048, "laplace", 77)
    c = CompressedMatrixV1.encode(w, 32, 8, 0, coder=1, outliers="auto").to_device()
    s_w = torch.full((512,), 0.01, device="cuda")
    x = torch.randn((3, 2048), device="cuda", requires_grad=True)
    y = mv.compressed_linear(x, c, s_w)
    gy = torch.randn_like(y)
    y.backward(gy)
    ref = (gy.double() * s_w.double()) @ torch.from_numpy(w).double().cuda()
    assert torch.allclose(x.grad.double(), ref, rtol=1e-5, atol=1e-5)
