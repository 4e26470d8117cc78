"""GPU checks of the backward dx = alpha * Wq^T (s_w * g) from the compressed streams:
against a dense dequantised FP64 reference, FP64 finite differences of the forward on a
256 x 512 slice, bitwise run-to-run determinism, and the autograd Function (STE)."""
import sys

import numpy as np
import pytest

import matvans_amd as mv
from matvans_amd.capi import MatvansError
from matvans_amd.mvx import CompressedMatrixV1
from oracle import oracle as O

from .helpers import gaussian_int8

pytestmark = pytest.mark.gpu


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


# q64r: 64 streams, one table per row; l32 / l64: the quad streams back to back (coder 2)
FORMATS = ["v0", "q32", "q64", "q128", "q64r", "l32", "l64"]


def encode(fmt, w):
    """The matrix on the device in one of the encodings with a backward: .bin v0, or the
    MVX1 quad coder with 32 / 64 / 128 streams and one table per matrix, or 64 streams and
    one table per row, or the quad streams back to back with 32 / 64 streams."""
    if fmt == "v0":
        return mv.CompressedMatrix.encode(w, n_threads=16).to_device()
    assert _fits(fmt, w.shape[1])
    return CompressedMatrixV1.encode(w, _streams(fmt), 8, int(fmt.endswith("r")),
                                     n_threads=16, coder=2 if fmt[0] == "l" else 1).to_device()


def dense_ref(w, s_w, g, alpha):
    """FP64 alpha[v] * W^T (s_w * g[v])."""
    return alpha[:, None] * ((g.astype(np.float64) * s_w[None, :].astype(np.float64)) @
                             w.astype(np.float64))


def _streams(fmt):
    return int(fmt[1:].rstrip("r"))


def _fits(fmt, cols):
    """MVX1 layouts need the column count to be a multiple of 16 * streams."""
    return fmt == "v0" or cols % (16 * _streams(fmt)) == 0


SHAPES = [(256, 512, 1), (100, 1024, 3), (4096, 14336, 1), (14336, 4096, 2), (1000, 2048, 8),
          (65, 2048, 5)]


@pytest.mark.parametrize("fmt,rows,cols,n_vec", [(f, *sh) for f in FORMATS for sh in SHAPES
                                                 if _fits(f, sh[1])])
def test_backward_vs_dense(cuda, fmt, rows, cols, n_vec):
    import torch
    note(f"backward {fmt} {rows}x{cols} n_vec={n_vec}")
    w, _ = gaussian_int8(rows, cols, 6.0, seed=rows + 3 * cols)
    cm = encode(fmt, w)
    rng = np.random.default_rng(7)
    g = rng.standard_normal((n_vec, rows)).astype(np.float32)
    s_w = (rng.random(rows) * 0.02 + 0.001).astype(np.float32)
    alpha = (rng.random(n_vec) + 0.5).astype(np.float32)
    dx = cm.backward(torch.from_numpy(g).cuda(), torch.from_numpy(s_w).cuda(),
                     torch.from_numpy(alpha).cuda())
    torch.cuda.synchronize()
    got = dx.cpu().numpy().astype(np.float64)
    ref = dense_ref(w, s_w, g, alpha.astype(np.float64))
    # norm-wise relative error (tolerance 1e-5) and an element-wise bound against the
    # sum of absolute terms
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    absum = alpha[:, None] * ((np.abs(g) * s_w[None, :]).astype(np.float64) @
                              np.abs(w).astype(np.float64))
    assert np.all(np.abs(got - ref) <= 1e-4 * absum + 1e-30)


@pytest.mark.parametrize("fmt", FORMATS)
def test_backward_deterministic(cuda, fmt):
    import torch
    note(f"backward determinism {fmt}")
    w, _ = gaussian_int8(2048, 4096, 6.0, seed=5)
    cm = encode(fmt, w)
    rng = np.random.default_rng(1)
    g = torch.from_numpy(rng.standard_normal((4, 2048)).astype(np.float32)).cuda()
    s_w = torch.from_numpy((rng.random(2048) + 0.1).astype(np.float32)).cuda()
    a = cm.backward(g, s_w).cpu().numpy()
    b = cm.backward(g, s_w).cpu().numpy()
    assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("fmt", ["v0", "q32"])
def test_finite_differences_slice(cuda, fmt):
    """dL/dxq of L(xq) = sum_r g_r * s_x * s_w[r] * (W xq)_r by FP64 central differences
    on a 256 x 512 slice equals the kernel's s_x * W^T (s_w * g)."""
    import torch
    note(f"finite differences 256x512 {fmt}")
    w, _ = gaussian_int8(256, 512, 6.0, seed=11)
    cm = encode(fmt, w)
    if fmt == "v0":
        wd = O.c_decode(O.Record(256, 512, 1.0, cm.cursors, cm.payload_size, cm.min_value, cm.G,
                                 cm.cdf_data, cm.ppf_data, cm.payload)).astype(np.float64)
    else:
        wd = w.astype(np.float64)   # the quad decode is bit-exact (test_gpu_v1)
    rng = np.random.default_rng(2)
    g = rng.standard_normal(256).astype(np.float32)
    s_w = (rng.random(256) * 0.1 + 0.01).astype(np.float32)
    s_x = np.float32(0.37)
    xq = rng.integers(-127, 128, 512).astype(np.float64)

    def loss(v):
        return float(np.sum(g.astype(np.float64) * float(s_x) * s_w.astype(np.float64) *
                            (wd @ v)))

    h = 0.5
    fd = np.array([(loss(xq + h * e) - loss(xq - h * e)) / (2 * h) for e in np.eye(512)])
    dx = cm.backward(torch.from_numpy(g).cuda(), torch.from_numpy(s_w).cuda(),
                     torch.tensor([s_x], device="cuda")).cpu().numpy().astype(np.float64)
    assert np.linalg.norm(dx - fd) <= 1e-5 * np.linalg.norm(fd)


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("batch", [1, 4, 8])
def test_autograd_function(cuda, fmt, batch):
    import torch
    note(f"autograd {fmt} batch={batch}")
    rows, cols = 1024, 2048
    w, _ = gaussian_int8(rows, cols, 6.0, seed=batch)
    cm = encode(fmt, w)
    rng = np.random.default_rng(batch)
    s_w = torch.from_numpy((rng.random(rows) * 0.02 + 0.001).astype(np.float32)).cuda()
    x = torch.from_numpy(rng.standard_normal((batch, cols)).astype(np.float32)).cuda()
    x.requires_grad_(True)
    y = mv.compressed_linear(x, cm, s_w)
    gy = torch.from_numpy(rng.standard_normal((batch, rows)).astype(np.float32)).cuda()
    (y * gy).sum().backward()
    torch.cuda.synchronize()
    # forward: s_x * s_w * (W @ xq) against FP64 with the same quantised xq
    xn = x.detach().cpu().numpy()
    s_x = np.abs(xn).max(axis=1) / np.float32(127)
    xq = np.clip(np.rint(xn / s_x[:, None]), -127, 127)
    yref = (s_x[:, None].astype(np.float64) * s_w.cpu().numpy()[None, :] *
            (xq @ w.astype(np.float64).T))
    assert np.all(np.abs(y.detach().cpu().numpy() - yref) <= 1e-6 * np.abs(yref) + 1e-30)
    # backward (straight-through): W^T (s_w * gy)
    gref = dense_ref(w, s_w.cpu().numpy(), gy.cpu().numpy(), np.ones(batch))
    got = x.grad.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(got - gref) <= 1e-5 * np.linalg.norm(gref)


@pytest.mark.parametrize("S,P,mode,coder", [(64, 8, 0, 0), (32, 12, 0, 0), (64, 8, 1, 0)])
def test_unsupported_encodings_are_refused(cuda, S, P, mode, coder):
    """The backward exists for .bin v0 and the quad coder; the single-symbol coders are
    refused with MATVANS_E_UNSUPPORTED before any launch (the output keeps its
    contents)."""
    import torch
    note(f"backward refuses S={S} P={P} mode={mode} coder={coder}")
    w, _ = gaussian_int8(64, 2048, 6.0, seed=9)
    cm = CompressedMatrixV1.encode(w, S, P, mode, coder=coder).to_device()
    g = torch.ones((1, 64), device="cuda")
    s_w = torch.ones(64, device="cuda")
    out = torch.full((1, 2048), float("nan"), device="cuda")
    with pytest.raises(MatvansError) as e:
        cm.backward(g, s_w, out=out)
    assert e.value.status == 8
    assert bool(torch.isnan(out).all())
