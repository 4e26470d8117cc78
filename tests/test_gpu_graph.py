"""HIP-graph capture of the single-matrix entry points.  Their launch path allocates
nothing and records no memset (buffers and workspaces come from the caller), so a graph
captured around them replays them: a two-layer .bin v0 chain (decode-matvec, device
absmax, requant, then the next matrix on the requantised vector), the quad-coder forward
and both backwards.  Every replay, with eager work and synchronisations in between,
reproduces the eager results bit for bit."""
import sys

import numpy as np
import pytest

import matvans_amd as mv
from matvans_amd import lib
from matvans_amd.mvx import CompressedMatrixV1
from oracle import oracle as O

from .helpers import gaussian_int8, uniform_vec

pytestmark = pytest.mark.gpu


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def test_single_matrix_entry_points_replay(cuda):
    import torch
    note("graph capture: v0 composite chain, quad forward, v0 and quad backward")
    rows, cols, rows2, n_vec = 512, 2048, 1024, 3
    m1 = gaussian_int8(rows, cols, 5.0, seed=81)[0]
    m2 = gaussian_int8(rows2, rows, 5.0, seed=82)[0]
    x = uniform_vec(cols, seed=83)
    cm1 = mv.CompressedMatrix.encode(m1).to_device()
    cm2 = mv.CompressedMatrix.encode(m2).to_device()
    q1 = CompressedMatrixV1.encode(m1, 32, 8, 0, coder=1).to_device()
    rng = np.random.default_rng(84)
    dev = "cuda"
    dx = torch.from_numpy(x).to(dev)
    g = torch.from_numpy(rng.standard_normal((n_vec, rows)).astype(np.float32)).to(dev)
    s_w = torch.from_numpy((rng.random(rows) * 0.02 + 0.001).astype(np.float32)).to(dev)
    alpha = torch.from_numpy((rng.random(n_vec) + 0.5).astype(np.float32)).to(dev)
    ws_bytes = int(lib().matvans_backward_workspace_size(rows, cols, n_vec))
    outs = dict(
        r32=torch.empty(rows, dtype=torch.int32, device=dev),
        r8=torch.empty(rows, dtype=torch.int8, device=dev),
        delta=torch.empty(1, dtype=torch.float32, device=dev),
        y2=torch.empty((1, rows2), dtype=torch.int32, device=dev),
        yq=torch.empty((1, rows), dtype=torch.int32, device=dev),
        bv0=torch.empty((n_vec, cols), dtype=torch.float32, device=dev),
        bq=torch.empty((n_vec, cols), dtype=torch.float32, device=dev))
    ws0 = torch.empty(max(1, ws_bytes // 4), dtype=torch.float32, device=dev)
    ws1 = torch.empty(max(1, ws_bytes // 4), dtype=torch.float32, device=dev)

    def body(s):
        cm1.decompress_and_mult(outs["r8"], outs["r32"], dx, outs["delta"], stream=s)
        cm2.multiply(outs["r8"], out=outs["y2"], stream=s)
        q1.multiply(dx, out=outs["yq"], stream=s)
        cm1.backward(g, s_w, alpha, out=outs["bv0"], workspace=ws0, stream=s)
        q1.backward(g, s_w, alpha, out=outs["bq"], workspace=ws1, stream=s)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        body(s)                       # eager, outside the capture
    s.synchronize()
    eager = {k: v.cpu().numpy().copy() for k, v in outs.items()}
    # the eager results themselves against the oracle
    want32 = O.matvec_i32(m1, x)
    want8, want_delta = O.requant_f32(want32)
    assert np.array_equal(eager["r32"], want32)
    assert np.array_equal(eager["r8"], want8) and eager["delta"][0] == want_delta
    assert np.array_equal(eager["y2"][0], O.matvec_i32(m2, want8))
    assert np.array_equal(eager["yq"][0], want32)
    gn, sn, an = g.cpu().numpy(), s_w.cpu().numpy(), alpha.cpu().numpy()
    ref = an[:, None].astype(np.float64) * ((gn.astype(np.float64) * sn.astype(np.float64))
                                            @ m1.astype(np.float64))
    for k in ("bv0", "bq"):   # norm-wise relative error, tolerance 1e-5
        assert np.linalg.norm(eager[k] - ref) <= 1e-5 * np.linalg.norm(ref), k

    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            body(s)
    torch.cuda.synchronize()
    for rep in range(3):
        for v in outs.values():
            v.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        t = torch.ones(1000, device=dev).sum()   # eager work between replays
        assert float(t) == 1000.0
        for k, v in outs.items():
            assert v.cpu().numpy().tobytes() == eager[k].tobytes(), (rep, k)
