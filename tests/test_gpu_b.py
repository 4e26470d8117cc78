"""GPU parity of the family-B fused decode-matvec (ans-warp-rows and ans-warp-stripes)
through the C ABI: the reference's own files (chain hashes, int32 sums), files encoded
by the library, 1..8 vectors, ragged tails and the int8 epilogue."""
import os
import sys

import ctypes

import numpy as np
import pytest

from matvans_amd.buffered import BufferedMatrices, read_safetensors
from matvans_amd.capi import MatvansError, call
from oracle import family_b as B

from .helpers import gaussian_int8, laplace_int8, uniform_vec

pytestmark = pytest.mark.gpu


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def want(m, vecs):
    return (m.astype(np.int64) @ np.asarray(vecs, np.int64).T).T.astype(np.int32)


@pytest.mark.parametrize("tag", ["256", "384_e5"])
@pytest.mark.parametrize("layout", ["rows", "stripes"])
def test_reference_files(cuda, golden_dir, tag, layout):
    import torch
    note(f"family-B file {layout} {tag}")
    raw = read_safetensors(os.path.join(golden_dir, f"b_raw_{tag}.safetensors"))
    bm = BufferedMatrices.read(os.path.join(golden_dir, f"b_{layout}_{tag}.safetensors"))
    bm.to_device()
    vecs = np.stack([raw["initial_vector"], uniform_vec(len(raw["initial_vector"]), seed=3)])
    dv = torch.from_numpy(vecs).cuda()
    for k, m in enumerate(raw["matrices"]):
        acc = bm.mat_vec(k, dv)
        torch.cuda.synchronize()
        assert np.array_equal(acc.cpu().numpy(), want(m, vecs))
    out = bm.run_chain()
    torch.cuda.synchronize()
    assert B.fxhash(out.cpu().numpy()) == int(raw["hashes"][-1])


CASES = [("rows", 32, 128, "g", 3.0), ("rows", 65, 384, "l", 2.0), ("rows", 300, 1024, "g", 8.0),
         ("stripes", 32, 4, "g", 3.0), ("stripes", 96, 260, "l", 2.0),
         ("stripes", 320, 1024, "g", 8.0)]


@pytest.mark.parametrize("layout,rows,cols,kind,p", CASES)
def test_library_encoded(cuda, layout, rows, cols, kind, p):
    import torch
    note(f"family-B {layout} {rows}x{cols} {kind}{p}")
    maker = gaussian_int8 if kind == "g" else laplace_int8
    m = maker(rows, cols, p, seed=rows + cols)[0]
    m2 = maker(rows, cols, p * 0.5, seed=7)[0]
    bm = BufferedMatrices.encode([m, m2], layout, scales=np.array([0.01, 0.003], np.float32))
    bm.to_device()
    for n_vec in (1, 3, 8):
        vecs = np.stack([uniform_vec(cols, seed=50 + i) for i in range(n_vec)])
        dv = torch.from_numpy(vecs).cuda()
        for k, mat in enumerate((m, m2)):
            acc = torch.empty((n_vec, rows), dtype=torch.int32, device="cuda")
            o8 = torch.empty((n_vec, rows), dtype=torch.int8, device="cuda")
            bm.mat_vec(k, dv, out8=o8, acc32=acc)
            torch.cuda.synchronize()
            w = want(mat, vecs)
            assert np.array_equal(acc.cpu().numpy(), w), (k, n_vec)
            assert np.array_equal(o8.cpu().numpy(), B.requant_b(w, bm.scales[k]))


def test_argument_errors(cuda):
    import torch
    note("family-B argument validation")
    m = gaussian_int8(32, 256, 3.0)[0]
    bm = BufferedMatrices.encode([m], "rows").to_device()
    v9 = torch.zeros((9, 256), dtype=torch.int8, device="cuda")
    with pytest.raises(ValueError, match="9 vectors"):   # the binding's check
        bm.mat_vec(0, v9)
    acc = torch.empty((9, 32), dtype=torch.int32, device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    with pytest.raises(MatvansError) as e:         # the C ABI's own check: 9 vectors
        call("matvans_b_rows_matvec", p(bm.d["compressed"]), bm.compressed.size,
             p(bm.d["offsets"][0]), bm.end_offset(0), p(bm.d["ppf"]), p(bm.d["cum"]),
             p(bm.d["prob"]), p(v9), 9, 256, 256, 32, 1.0, None, p(acc), None)
    assert e.value.status == 1
    bm.shapes[0] = (32, 200)                       # rows layout needs cols % 128 == 0
    with pytest.raises(MatvansError) as e:
        bm.mat_vec(0, torch.zeros((1, 200), dtype=torch.int8, device="cuda"))
    assert e.value.status == 2
