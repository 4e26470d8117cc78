"""GPU parity of the MVX1 fused decode-matvec (64 streams per row, 12-bit probabilities,
one table per row, the quad coder in both layouts) against the brute-force integer product and the oracle decoder.
Every launch goes through the C ABI (matvans_v1_decompress_and_multiply)."""
import ctypes
import sys

import numpy as np
import pytest

from matvans_amd.capi import MatvansError, call
from matvans_amd.mvx import CompressedMatrixV1
from oracle import ext as E

from .helpers import gaussian_int8, gaussian_vec, laplace_int8, uniform_vec

pytestmark = pytest.mark.gpu

VARIANTS = [(32, 8, 0, 0), (64, 8, 0, 0), (128, 8, 0, 0), (32, 12, 0, 0), (64, 12, 0, 0),
            (128, 12, 0, 0), (64, 8, 1, 0), (32, 8, 0, 1), (64, 8, 0, 1), (128, 8, 0, 1),
            (64, 8, 1, 1), (32, 8, 0, 2), (64, 8, 0, 2), (128, 8, 0, 2)]


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def gpu_matvec(c, vecs):
    import torch
    v = torch.from_numpy(np.ascontiguousarray(vecs)).cuda()
    out = c.multiply(v)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def want(m, vecs):
    return (m.astype(np.int64) @ np.asarray(vecs, np.int64).T).T.astype(np.int32)


SMALL = [(1, 2048, "g", 6.0), (3, 1024, "g", 0.3), (7, 3072, "l", 4.0), (65, 2048, "g", 16.0),
         (129, 4096, "l", 4.0), (333, 1024, "g", 60.0)]


@pytest.mark.parametrize("S,P,mode,coder,rows,cols,kind,p",
                         [(*v, *sh) for v in VARIANTS for sh in SMALL
                          if sh[1] % (16 * v[0]) == 0])   # cols a multiple of 16 * streams
def test_shapes(cuda, S, P, mode, coder, rows, cols, kind, p):
    note(f"v1 S={S} P={P} mode={mode} coder={coder} {rows}x{cols} {kind}{p}")
    maker = gaussian_int8 if kind == "g" else laplace_int8
    m, _ = maker(rows, cols, p, seed=rows + 3 * cols)
    c = CompressedMatrixV1.encode(m, S, P, mode, coder=coder).to_device()
    for n_vec in (1, 3, 8):
        vecs = np.stack([uniform_vec(cols, seed=200 + i) for i in range(n_vec)])
        got = gpu_matvec(c, vecs)
        assert np.array_equal(got, want(m, vecs)), (n_vec, np.argwhere(got != want(m, vecs))[:4])


@pytest.mark.parametrize("S,P,mode,coder", VARIANTS)
def test_edge_alphabets(cuda, S, P, mode, coder):
    """Constant rows, a near-constant matrix (few renormalisations) and a uniform
    wide-alphabet matrix (a renormalisation at almost every step)."""
    note(f"v1 edge alphabets S={S} P={P} mode={mode} coder={coder}")
    near = np.zeros((9, 2048), np.int8)
    near[1, 5] = 2
    near[4] = -127
    near[7] = 127
    wide = np.random.default_rng(1).integers(-128, 128, (6, 2048)).astype(np.int8)
    for m in (near, wide, np.full((4, 2048), -3, np.int8)):
        c = CompressedMatrixV1.encode(m, S, P, mode, coder=coder).to_device()
        vecs = np.stack([uniform_vec(2048, seed=1), np.full(2048, -128, np.int8)])
        got = gpu_matvec(c, vecs)
        assert np.array_equal(got, want(m, vecs))
        r = E.encode_record(m, S, P, mode, coder=coder)   # oracle decode of the same bytes
        assert np.array_equal(r.decode(), m)


LLM = [(4096, 14336, "l", 4.0), (14336, 4096, "l", 4.0), (4096, 4096, "g", 6.0)]


@pytest.mark.parametrize("S,P,mode,coder", [(128, 8, 0, 0), (64, 8, 0, 0), (32, 8, 0, 0),
                                            (64, 12, 0, 0), (64, 8, 1, 0), (32, 12, 0, 0),
                                            (32, 8, 0, 1), (64, 8, 0, 1), (128, 8, 0, 1),
                                            (64, 8, 1, 1), (32, 8, 0, 2), (64, 8, 0, 2)])
@pytest.mark.parametrize("rows,cols,kind,p", LLM)
def test_llm_shapes(cuda, S, P, mode, coder, rows, cols, kind, p):
    note(f"v1 LLM S={S} P={P} mode={mode} coder={coder} {rows}x{cols} {kind}{p}")
    maker = gaussian_int8 if kind == "g" else laplace_int8
    m, _ = maker(rows, cols, p, seed=rows + cols)
    c = CompressedMatrixV1.encode(m, S, P, mode, n_threads=16, coder=coder).to_device()
    vecs = np.stack([gaussian_vec(cols, seed=3), uniform_vec(cols, seed=4)])
    assert np.array_equal(gpu_matvec(c, vecs), want(m, vecs))


def test_argument_errors(cuda):
    import torch
    note("v1 argument validation")
    m, _ = gaussian_int8(8, 2048, 6.0)
    c = CompressedMatrixV1.encode(m, 64, 8, 0).to_device()
    v = torch.zeros((9, 2048), dtype=torch.int8, device="cuda")
    with pytest.raises(ValueError, match="9 vectors"):   # the binding's check
        c.multiply(v)
    out9 = torch.empty((9, 8), dtype=torch.int32, device="cuda")
    with pytest.raises(MatvansError) as e:      # the C ABI's own check: 9 vectors
        call("matvans_v1_decompress_and_multiply", ctypes.byref(c.dev_struct()),
             ctypes.c_void_p(out9.data_ptr()), ctypes.c_void_p(v.data_ptr()), 9, 2048, None)
    assert e.value.status == 1
    s = c.dev_struct()
    s.table_mode = 1
    s.precision = 12                           # 12-bit per-row tables: not a format
    out = torch.empty(8, dtype=torch.int32, device="cuda")
    with pytest.raises(MatvansError) as e:
        call("matvans_v1_decompress_and_multiply", ctypes.byref(s), ctypes.c_void_p(out.data_ptr()),
             ctypes.c_void_p(v.data_ptr()), 1, 2048, None)
    assert e.value.status == 8
    s = c.dev_struct()
    s.payload = s.payload + 2                  # misaligned payload
    with pytest.raises(MatvansError) as e:
        call("matvans_v1_decompress_and_multiply", ctypes.byref(s), ctypes.c_void_p(out.data_ptr()),
             ctypes.c_void_p(v.data_ptr()), 1, 2048, None)
    assert e.value.status == 3
