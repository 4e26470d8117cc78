"""The batched workloads at their full sizes, bit-exact against the integer product.

* Decoder block: q 4096x4096, k and v 1024x4096, o 4096x4096, gate and up 14336x4096,
  down 4096x14336, int8 entries of sigma 6, with 1, 4 and 8 activation vectors, in one
  persistent launch per step, for the .bin v0 encoding and the quad coder with 32 and 64
  streams per row.
* Uneven batch: 32 matrices with rows ~ U[256, 16384], cols ~ U{512, 768, ..., 8192},
  sigma in {2, 4, 8, 16}, row and stripe layouts mixed (rows: quad coder with 32 streams,
  columns zero-padded to a multiple of 512 when needed; stripes: the family-B stripe
  layout), 1..8 vectors each, in one persistent launch.
* The backward at eight vectors at both LLM shapes, against dense FP64.

The integer product is computed in float64 on the host (exact: |sums| < 2^53)."""
import sys

import numpy as np
import pytest

import matvans_amd as mv
from matvans_amd.batch import Batch
from matvans_amd.buffered import BufferedMatrices
from matvans_amd.mvx import CompressedMatrixV1

from .helpers import BLOCK_SHAPES, act_vectors, pad_cols, sample_int8, uneven_specs

pytestmark = pytest.mark.gpu


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def exact(m, v):
    return (np.asarray(v, np.float64) @ np.asarray(m, np.float64).T).astype(np.int32)


@pytest.fixture(scope="module")
def block():
    rng = np.random.default_rng(20250311)
    mats = [sample_int8(rng, (r, c), 6.0) for _, r, c in BLOCK_SHAPES]
    xs = [act_vectors(rng, 8, c) for _, _, c in BLOCK_SHAPES]
    return mats, xs


_ENCODED = {}


def block_encoded(fmt, mats):
    if fmt not in _ENCODED:
        _ENCODED.clear()   # one encoding resident at a time
        if fmt == "v0":
            _ENCODED[fmt] = [mv.CompressedMatrix.encode(m).to_device() for m in mats]
        else:
            coder = 2 if fmt[0] == "l" else 1
            _ENCODED[fmt] = [CompressedMatrixV1.encode(m, int(fmt[1:]), 8, 0, n_threads=16,
                                                       coder=coder).to_device() for m in mats]
    return _ENCODED[fmt]


@pytest.mark.parametrize("fmt", ["v0", "q32", "q64", "l32"])
@pytest.mark.parametrize("n_vec", [1, 4, 8])
def test_decoder_block_full_size(cuda, block, fmt, n_vec):
    import torch
    mats, xs = block
    note(f"full decoder block {fmt} n_vec={n_vec}")
    cms = block_encoded(fmt, mats)
    dv = [torch.from_numpy(x[:n_vec].copy()).cuda() for x in xs]
    b = Batch(list(zip(cms, dv)))
    for rep in range(2):            # the work queue resets itself between launches
        for o in b.outputs:
            o.fill_(-1)
        b.launch()
        torch.cuda.synchronize()
        for (name, _, _), m, x, o in zip(BLOCK_SHAPES, mats, xs, b.outputs):
            assert np.array_equal(o.cpu().numpy(), exact(m, x[:n_vec])), (name, rep)
    b.close()


def test_uneven_batch_full_size(cuda):
    """32 matrices in one persistent launch (1.2 G entries), default plan and a second
    configuration."""
    import torch
    specs = uneven_specs()
    note(f"uneven batch, {len(specs)} matrices")
    rng = np.random.default_rng(20250313)
    entries, checks = [], []
    stripes = [i for i, s in enumerate(specs) if s[3] == "stripes"]
    mats = [sample_int8(rng, (r, c), s) for r, c, s, _, _ in specs]
    xs = [act_vectors(rng, nv, c) for (_, c, _, _, nv) in specs]
    bm = BufferedMatrices.encode([mats[i] for i in stripes], "stripes").to_device()
    where = {i: k for k, i in enumerate(stripes)}
    for i, (rows, cols, sigma, layout, nv) in enumerate(specs):
        if layout == "stripes":
            assert np.abs(mats[i].astype(np.int32)).max() < 64   # 7-bit model range
            entries.append(((bm, where[i]), torch.from_numpy(xs[i]).cuda()))
        else:
            mp = pad_cols(mats[i], 512)
            xp = np.zeros((nv, mp.shape[1]), np.int8)
            xp[:, :cols] = xs[i]
            c = CompressedMatrixV1.encode(mp, 32, 8, 0, coder=1).to_device()
            entries.append((c, torch.from_numpy(xp).cuda()))
        checks.append(exact(mats[i], xs[i]))
    assert 0 < len(stripes) < len(specs)
    assert any(s[1] % 512 for s in specs)    # some row-layout matrices need the padding
    b = Batch(entries)
    for cfg in [None, (2, 1, 4, 0)]:
        if cfg is not None:
            b.configure(*cfg)
        for o in b.outputs:
            o.fill_(-1)
        b.launch()
        torch.cuda.synchronize()
        for i, (o, w) in enumerate(zip(b.outputs, checks)):
            assert np.array_equal(o.cpu().numpy(), w), (i, specs[i], cfg)
    b.close()


@pytest.mark.parametrize("fmt", ["v0", "q32", "q64", "l32", "l64"])
@pytest.mark.parametrize("rows,cols", [(4096, 14336), (14336, 4096)])
def test_backward_eight_vectors_llm_shapes(cuda, fmt, rows, cols):
    """dx = alpha * W^T (s_w * g) for 8 gradient vectors from the compressed streams,
    against dense FP64 (norm-wise relative error <= 1e-5)."""
    import torch
    note(f"backward n_vec=8 {fmt} {rows}x{cols}")
    rng = np.random.default_rng(rows + cols)
    w = sample_int8(rng, (rows, cols), 6.0)
    cm = (mv.CompressedMatrix.encode(w) if fmt == "v0"
          else CompressedMatrixV1.encode(w, int(fmt[1:]), 8, 0, n_threads=16,
                                         coder=2 if fmt[0] == "l" else 1)).to_device()
    g = rng.standard_normal((8, rows)).astype(np.float32)
    s_w = (rng.random(rows) * 0.02 + 0.001).astype(np.float32)
    alpha = (rng.random(8) + 0.5).astype(np.float32)
    dx = cm.backward(torch.from_numpy(g).cuda(), torch.from_numpy(s_w).cuda(),
                     torch.from_numpy(alpha).cuda())
    torch.cuda.synchronize()
    got = dx.cpu().numpy().astype(np.float64)
    ref = alpha.astype(np.float64)[:, None] * ((g.astype(np.float64) * s_w[None, :]) @
                                               w.astype(np.float64))
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)


@pytest.mark.parametrize("fmt", ["v0", "q32", "q64", "l32"])
@pytest.mark.parametrize("n_vec", [1, 3])
def test_backward_one_matrix_equals_batch_at_14336_rows(cuda, fmt, n_vec):
    """At 14336 rows the one-matrix backward's 8-wave grid would overflow the GPU, so it
    runs units of 4 waves; the batched backward runs 8-wave units (and, at 3 vectors, its
    4-vector instance).  Every wave computes the same: dx bitwise equal."""
    import torch
    from matvans_amd.batch import BatchBackward
    note(f"backward one-matrix vs batch 14336x4096 {fmt} n_vec={n_vec}")
    rng = np.random.default_rng(31 + n_vec)
    w = sample_int8(rng, (14336, 4096), 6.0)
    cm = (mv.CompressedMatrix.encode(w) if fmt == "v0"
          else CompressedMatrixV1.encode(w, int(fmt[1:]), 8, 0, n_threads=16,
                                         coder=2 if fmt[0] == "l" else 1)).to_device()
    g = torch.from_numpy(rng.standard_normal((n_vec, 14336)).astype(np.float32)).cuda()
    s_w = torch.from_numpy((rng.random(14336) * 0.02 + 0.001).astype(np.float32)).cuda()
    alpha = torch.from_numpy((rng.random(n_vec) + 0.5).astype(np.float32)).cuda()
    single = cm.backward(g, s_w, alpha).cpu().numpy()
    b = BatchBackward([(cm, g, s_w, alpha)])
    b.launch()
    torch.cuda.synchronize()
    assert b.outputs[0].cpu().numpy().tobytes() == single.tobytes()
    b.close()
