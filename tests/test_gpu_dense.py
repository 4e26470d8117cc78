"""GPU parity of the comparators and epilogues: int8 sdot4 matvec (bit-exact),
fp32 GEMV (vs FP64), absmax/requant (bit-exact vs the fp32 oracle epilogue), the fp32
chain's normalize-and-round (bit-exact vs the oracle) and the scaled fp32 output (1e-6
relative vs FP64)."""
import sys

import numpy as np
import pytest

import matvans_amd as mv
from oracle import oracle as O

from .helpers import gaussian_int8, uniform_vec

pytestmark = pytest.mark.gpu


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


# 4-wave blocks, one row per wave with the vectors in LDS (small cases, 4096 x 14336 x 1)
# or two from 8192 rows (8193: a last wave with one row); persistent 16-wave blocks for
# 16-160 KiB of vectors (7 x 8192 x 5, 100 x 20000 x 3, 14336 x 4096 x 8, 4097 x 14336 x
# 8: more rows than the grid's waves); two rows per wave from the caches beyond 160 KiB
# (5 x 24576 x 7)
@pytest.mark.parametrize("rows,cols,n_vec", [(1, 16, 1), (5, 48, 2), (37, 4096, 1),
                                             (4096, 14336, 1), (14336, 4096, 8),
                                             (100, 20000, 3), (8193, 512, 2), (7, 8192, 5),
                                             (4097, 14336, 8), (5, 24576, 7)])
def test_int8_matvec(cuda, rows, cols, n_vec):
    import torch
    note(f"int8 {rows}x{cols} nvec={n_vec}")
    m, _ = gaussian_int8(rows, cols, 20.0, seed=rows)
    vecs = np.stack([uniform_vec(cols, seed=i) for i in range(n_vec)])
    got = mv.int8_matvec(torch.from_numpy(m).cuda(), torch.from_numpy(vecs).cuda())
    torch.cuda.synchronize()
    want = (m.astype(np.int64) @ vecs.astype(np.int64).T).T
    assert np.array_equal(got.cpu().numpy(), want.astype(np.int32))


@pytest.mark.parametrize("rows,cols", [(3, 4), (64, 1024), (4096, 14336), (17, 20000)])
def test_f32_gemv(cuda, rows, cols):
    import torch
    note(f"f32 gemv {rows}x{cols}")
    rng = np.random.default_rng(rows)
    a = rng.standard_normal((rows, cols)).astype(np.float32)
    x = rng.standard_normal(cols).astype(np.float32)
    got = mv.f32_gemv(torch.from_numpy(a).cuda(), torch.from_numpy(x).cuda(), alpha=0.5)
    torch.cuda.synchronize()
    want = 0.5 * (a.astype(np.float64) @ x.astype(np.float64))
    scale = 0.5 * (np.abs(a).astype(np.float64) @ np.abs(x).astype(np.float64))
    assert np.all(np.abs(got.cpu().numpy() - want) <= 1e-5 * scale)


def edge_max_cases(rng):
    """the extreme value in the last, partial stride of the kernel's loop (n past a
    multiple of the block) and a negative extreme"""
    out = []
    for n, at, sign in [(4097, 4096, 1), (5001, 5000, -1), (1025, 1, 1)]:
        a = rng.integers(-1000, 1000, n).astype(np.int32)
        a[at] = sign * 123457
        out.append(a)
    return out


@pytest.mark.parametrize("clamp", [False, True])
def test_absmax_requant(cuda, clamp):
    import torch
    note(f"absmax requant clamp={clamp}")
    rng = np.random.default_rng(1)
    for acc in [rng.integers(-2**20, 2**20, 4096).astype(np.int32),
                np.array([254, 1, 3, 5, -5, -3], np.int32),            # ties
                rng.integers(-2**30, 2**30, 14336).astype(np.int32),    # > 2^24
                np.zeros(7, np.int32)] + edge_max_cases(rng):
        out8, d = mv.absmax_requant(torch.from_numpy(acc).cuda(), clamp=clamp)
        torch.cuda.synchronize()
        ref8, refd = O.requant_f32(acc, clamp=clamp)
        assert np.array_equal(out8.cpu().numpy(), ref8)
        if acc.any():
            assert d.cpu().numpy()[0] == refd
        else:
            assert np.array_equal(out8.cpu().numpy(), np.zeros(7, np.int8))


def test_f32_normalize_round(cuda):
    import torch
    note("f32 normalize and round")
    rng = np.random.default_rng(3)
    cases = [(rng.standard_normal(n) * sc).astype(np.float32)
             for n, sc in [(1, 1.0), (1000, 3e4), (4096, 1e-3), (14336, 7.0), (100003, 1e6)]]
    cases.append(np.array([127.0, 0.5, -0.5, 1.5, 2.5, -2.5, 0.49999997, -126.5], np.float32))
    for y in cases:
        d_y = torch.from_numpy(y.copy()).cuda()
        _, d = mv.f32_normalize_round(d_y)
        torch.cuda.synchronize()
        ref, refd = O.normalize_round_f32(y)
        assert d.cpu().numpy()[0] == refd
        assert np.array_equal(d_y.cpu().numpy(), ref)
    # all zeros: delta 0 and NaN, as in the reference
    d_y = torch.zeros(5, device="cuda")
    _, d = mv.f32_normalize_round(d_y)
    torch.cuda.synchronize()
    assert d.cpu().numpy()[0] == 0 and bool(torch.isnan(d_y).all())


def test_scale_output(cuda):
    import torch
    note("scale output")
    rng = np.random.default_rng(2)
    rows, n_vec = 4096, 3
    acc = rng.integers(-2**28, 2**28, (n_vec, rows)).astype(np.int32)
    s_w = (rng.random(rows) + 0.01).astype(np.float32)
    s_x = (rng.random(n_vec) + 0.01).astype(np.float32)
    y = mv.scale_output(torch.from_numpy(acc).cuda(), torch.from_numpy(s_w).cuda(),
                        torch.from_numpy(s_x).cuda())
    torch.cuda.synchronize()
    ref = s_x.astype(np.float64)[:, None] * s_w.astype(np.float64)[None, :] * acc
    rel = np.abs(y.cpu().numpy() - ref) / np.maximum(np.abs(ref), 1e-30)
    assert rel.max() <= 1e-6
