"""GPU tests of the persistent batched decode-matvec (matvans_batch_*): bit-exact against
the integer product and the one-matrix launches, for a decoder block and an uneven mixed
batch; self-resetting work queue across launches; HIP-graph replays; autotuner."""
import sys

import numpy as np
import pytest

import matvans_amd as mv
from matvans_amd.batch import Batch
from matvans_amd.buffered import BufferedMatrices
from matvans_amd.capi import MatvansError
from matvans_amd.mvx import CompressedMatrixV1
from oracle import family_b as B

from .helpers import gaussian_int8, uniform_vec

pytestmark = pytest.mark.gpu


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def want(m, vecs):
    return (m.astype(np.int64) @ np.asarray(vecs, np.int64).T).T.astype(np.int32)


def vecs_for(cols, n, seed):
    return np.stack([uniform_vec(cols, seed=seed + i) for i in range(n)])


# decoder-block shapes at 1/2 scale (4096 x 14336 block): q, k, v, o, gate, up, down
BLOCK = [(2048, 2048), (512, 2048), (512, 2048), (2048, 2048), (7168, 2048), (7168, 2048),
         (2048, 7168)]


def make(fmt, m):
    """v0, q32 / q64 (quad coder), l32 / l64 (quad streams back to back, coder 2)."""
    if fmt == "v0":
        return mv.CompressedMatrix.encode(m).to_device()
    return CompressedMatrixV1.encode(m, int(fmt[1:]), 8, 0,
                                     coder=2 if fmt[0] == "l" else 1).to_device()


@pytest.mark.parametrize("fmt", ["v0", "q32", "q64", "l32", "l64"])
@pytest.mark.parametrize("n_vec", [1, 4, 8])
def test_decoder_block(cuda, fmt, n_vec):
    import torch
    note(f"batch decoder block {fmt} n_vec={n_vec}")
    mats = [gaussian_int8(r, c, 6.0, seed=r + 7 * c + i)[0] for i, (r, c) in enumerate(BLOCK)]
    cms = [make(fmt, m) for m in mats]
    vs = [vecs_for(m.shape[1], n_vec, 10 * i) for i, m in enumerate(mats)]
    b = Batch([(c, torch.from_numpy(v).cuda()) for c, v in zip(cms, vs)])
    for rep in range(3):           # the queue resets itself between launches
        for o in b.outputs:
            o.zero_()
        b.launch()
        torch.cuda.synchronize()
        for m, v, o in zip(mats, vs, b.outputs):
            assert np.array_equal(o.cpu().numpy(), want(m, v)), rep


def test_mixed_uneven_batch(cuda):
    """.bin v0, quad 32 / 64, family-B rows and stripes, uneven shapes, 1..8 vectors."""
    import torch
    note("batch mixed uneven")
    rng = np.random.default_rng(3)
    entries, checks = [], []
    specs = [("v0", 300, 1024), ("q32", 77, 2048), ("q64", 129, 1024), ("brows", 90, 384),
             ("bstripes", 96, 260), ("v0", 5, 512), ("q64", 1000, 2048), ("brows", 33, 128),
             ("bstripes", 32, 1024), ("q128", 131, 2048), ("q128", 1, 4096)]
    for i, (fmt, rows, cols) in enumerate(specs):
        sigma = [2.0, 4.0, 8.0, 16.0][i % 4]
        m = gaussian_int8(rows, cols, sigma if fmt[0] != "b" else min(sigma, 8.0), seed=i)[0]
        n_vec = int(rng.integers(1, 9))
        v = vecs_for(cols, n_vec, 100 * i)
        dv = torch.from_numpy(v).cuda()
        if fmt[0] == "b":
            bm = BufferedMatrices.encode([m], "rows" if fmt == "brows" else "stripes",
                                         scales=np.array([0.01], np.float32)).to_device()
            entries.append(((bm, 0), dv))
            single = bm.mat_vec(0, dv)
        else:
            c = make(fmt, m)
            entries.append((c, dv))
            single = c.multiply(dv)
        checks.append((m, v, single, fmt))
    b = Batch(entries)
    b.launch()
    torch.cuda.synchronize()
    for (m, v, single, fmt), o, o8 in zip(checks, b.outputs, b.out8):
        got = o.cpu().numpy()
        assert np.array_equal(got, want(m, v)), fmt
        assert np.array_equal(got, single.cpu().numpy()), fmt
        if o8 is not None:
            assert np.array_equal(o8.cpu().numpy(), B.requant_b(want(m, v), 0.01))
    for cfg in [(1, 1, 4), (4, 2, 8), (2, 0, 2), (1, 2, 8, 1), (4, 1, 3, 1)]:
        b.configure(*cfg)
        for o in b.outputs:
            o.zero_()
        b.launch()
        torch.cuda.synchronize()
        for (m, v, _s, fmt), o in zip(checks, b.outputs):
            assert np.array_equal(o.cpu().numpy(), want(m, v)), (cfg, fmt)


def test_int8_epilogue_every_format(cuda):
    """The next layer's int8 input computed in the launch, for .bin v0, quad and family-B
    items: out8 = clamp(rint(float(acc) * scale)) (the reference's family-B epilogue,
    restated by the oracle), with and without the int32 output, and a two-layer chain
    whose second plan reads the first plan's out8 as its vectors."""
    import torch
    note("batch int8 epilogue, every format")
    specs = [("v0", 300, 1024, 0.003), ("q32", 77, 2048, 0.0125), ("q64", 129, 1024, 0.05),
             ("brows", 90, 384, 0.02), ("q32", 1024, 4096, 0.004)]
    entries, checks = [], []
    for i, (fmt, rows, cols, scale) in enumerate(specs):
        m = gaussian_int8(rows, cols, 6.0, seed=50 + i)[0]
        n_vec = 1 + i % 3
        v = vecs_for(cols, n_vec, 300 + 10 * i)
        dv = torch.from_numpy(v).cuda()
        want32 = fmt != "q64"          # the q64 item writes only its int8 output
        if fmt == "brows":
            bm = BufferedMatrices.encode([m], "rows", scales=np.array([0.01], np.float32))
            entries.append(((bm.to_device(), 0), dv, scale))   # overrides the file's scale
        else:
            entries.append((make(fmt, m), dv, scale, want32))
        checks.append((m, v, scale, want32))
    b = Batch(entries)
    for rep in range(2):
        b.launch()
        torch.cuda.synchronize()
        for (m, v, scale, want32), o, o8 in zip(checks, b.outputs, b.out8):
            acc = want(m, v)
            assert np.array_equal(o8.cpu().numpy(), B.requant_b(acc, scale)), (rep, scale)
            if want32:
                assert np.array_equal(o.cpu().numpy(), acc)
            else:
                assert o is None
    # two layers: 1024 int8 outputs of the last item feed a 512 x 1024 quad matrix
    m2 = gaussian_int8(512, 1024, 6.0, seed=99)[0]
    first = b.out8[-1]                 # [n_vec, 1024]
    b2 = Batch([(make("q32", m2), first, 0.01)])
    b.launch()
    b2.launch()
    torch.cuda.synchronize()
    x1 = B.requant_b(want(checks[-1][0], checks[-1][1]), checks[-1][2])
    assert np.array_equal(b2.outputs[0].cpu().numpy(), want(m2, x1))
    assert np.array_equal(b2.out8[0].cpu().numpy(), B.requant_b(want(m2, x1), 0.01))
    # the same chain captured in one graph and replayed
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            b.launch(stream=s)
            b2.launch(stream=s)
    for rep in range(3):
        b2.outputs[0].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(b2.outputs[0].cpu().numpy(), want(m2, x1)), rep


def test_entry_without_output_is_refused(cuda):
    import torch
    note("batch entry without output")
    m = gaussian_int8(64, 1024, 6.0, seed=1)[0]
    dv = torch.from_numpy(vecs_for(1024, 1, 1)).cuda()
    with pytest.raises(ValueError):
        Batch([(make("q32", m), dv, None, False)])


def test_graph_replays_and_autotune(cuda):
    import torch
    note("batch graph replays + autotune")
    mats = [gaussian_int8(r, c, 6.0, seed=i)[0] for i, (r, c) in enumerate(BLOCK)]
    cms = [make("q32", m) for m in mats]
    vs = [vecs_for(m.shape[1], 2, i) for i, m in enumerate(mats)]
    dvs = [torch.from_numpy(v).cuda() for v in vs]
    b = Batch(list(zip(cms, dvs)))
    cfg, ms = b.autotune(reps=2)
    assert cfg == b.config and ms > 0
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        b.launch(stream=s)        # warm-up outside the capture
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            b.launch(stream=s)
    torch.cuda.synchronize()
    for rep in range(3):
        for o in b.outputs:
            o.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        x = torch.ones(1000, device="cuda").sum()   # eager work between replays
        assert float(x) == 1000.0
        for m, v, o in zip(mats, vs, b.outputs):
            assert np.array_equal(o.cpu().numpy(), want(m, v)), rep


def test_argument_errors(cuda):
    import torch
    note("batch argument validation")
    m = gaussian_int8(8, 2048, 6.0)[0]
    c = make("q32", m)
    v9 = torch.zeros((9, 2048), dtype=torch.int8, device="cuda")
    with pytest.raises(ValueError, match="9 vectors"):   # the binding's check
        Batch([(c, v9)])
    # the C ABI's own check on the same item
    import ctypes
    from matvans_amd.capi import BatchItem, call
    dst = torch.empty((9, 8), dtype=torch.int32, device="cuda")
    it = (BatchItem * 1)()
    it[0].format, it[0].rows, it[0].cols, it[0].n_vec = 1, 8, 2048, 9
    it[0].vec_stride, it[0].vecs, it[0].dst = 2048, v9.data_ptr(), dst.data_ptr()
    it[0].cursors, it[0].payload, it[0].payload_size = (c.d_cursors.data_ptr(),
                                                        c.d_payload.data_ptr(), c.payload_size)
    it[0].streams, it[0].lut = 32, c.d_lut.data_ptr()
    h = ctypes.c_void_p()
    with pytest.raises(MatvansError) as e:
        call("matvans_batch_create", it, 1, ctypes.byref(h))
    assert e.value.status == 1


# ---- batched backward: bitwise equal to each matrix's own backward

def bwd_inputs(m, n_vec, seed):
    import torch
    rng = np.random.default_rng(seed)
    g = torch.from_numpy(rng.standard_normal((n_vec, m.shape[0])).astype(np.float32)).cuda()
    s_w = torch.from_numpy((rng.random(m.shape[0]) * 0.02 + 0.001).astype(np.float32)).cuda()
    alpha = torch.from_numpy((rng.random(n_vec) + 0.5).astype(np.float32)).cuda()
    return g, s_w, alpha


@pytest.mark.parametrize("fmt", ["v0", "q32", "q64", "l32", "l64"])
@pytest.mark.parametrize("n_vec", [1, 8])
def test_backward_batch_block(cuda, fmt, n_vec):
    import torch
    from matvans_amd.batch import BatchBackward
    note(f"backward batch decoder block {fmt} n_vec={n_vec}")
    mats = [gaussian_int8(r, c, 6.0, seed=r + 7 * c + i)[0] for i, (r, c) in enumerate(BLOCK)]
    cms = [make(fmt, m) for m in mats]
    ins = [bwd_inputs(m, n_vec, i) for i, m in enumerate(mats)]
    singles = [c.backward(g, s, a).cpu().numpy() for c, (g, s, a) in zip(cms, ins)]
    b = BatchBackward([(c, g, s, a) for c, (g, s, a) in zip(cms, ins)])
    for cfg in [None, (1, 2, 4, 0), (4, 0, 2, 1), (2, 1, 8, 0)]:
        if cfg:
            b.configure(*cfg)
        for rep in range(2):
            for o in b.outputs:
                o.zero_()
            b.launch()
            torch.cuda.synchronize()
            for o, want in zip(b.outputs, singles):
                assert o.cpu().numpy().tobytes() == want.tobytes(), (cfg, rep)


def test_backward_batch_mixed_and_graph(cuda):
    import torch
    from matvans_amd.batch import BatchBackward
    note("backward batch mixed uneven + graph replays")
    rng = np.random.default_rng(11)
    specs = [("v0", 65, 1024), ("q32", 1000, 2048), ("q64", 100, 1024), ("v0", 300, 512),
             ("q32", 7, 512), ("q64", 2049, 3072)]
    entries, singles, dense = [], [], []
    for i, (fmt, r, c) in enumerate(specs):
        m = gaussian_int8(r, c, [2.0, 6.0, 16.0][i % 3], seed=50 + i)[0]
        cm = make(fmt, m)
        n_vec = int(rng.integers(1, 9))
        g, s_w, alpha = bwd_inputs(m, n_vec, 100 + i)
        entries.append((cm, g, s_w, alpha))
        singles.append(cm.backward(g, s_w, alpha).cpu().numpy())
        ref = alpha.cpu().numpy()[:, None].astype(np.float64) * (
            (g.cpu().numpy().astype(np.float64) * s_w.cpu().numpy()[None, :]) @ m.astype(np.float64))
        dense.append(ref)
    b = BatchBackward(entries)
    b.autotune(reps=2)
    b.launch()
    torch.cuda.synchronize()
    for o, want, ref in zip(b.outputs, singles, dense):
        got = o.cpu().numpy()
        assert got.tobytes() == want.tobytes()
        assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    gph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        b.launch(stream=s)
        s.synchronize()
        with torch.cuda.graph(gph, stream=s):
            b.launch(stream=s)
    torch.cuda.synchronize()
    for rep in range(3):
        for o in b.outputs:
            o.zero_()
        torch.cuda.synchronize()
        gph.replay()
        torch.cuda.synchronize()
        assert float(torch.ones(100, device="cuda").sum()) == 100.0   # eager work in between
        for o, want in zip(b.outputs, singles):
            assert o.cpu().numpy().tobytes() == want.tobytes(), rep


def test_tune_picks_an_exact_plan(cuda):
    """The streams-per-row / plan tuner returns a plan whose outputs are exact."""
    import torch
    from matvans_amd.batch import tune
    note("batch tune over encodings")
    mats = [gaussian_int8(r, c, 6.0, seed=7 + i)[0] for i, (r, c) in enumerate(BLOCK[:4])]
    vs = [vecs_for(m.shape[1], 2, 40 + i) for i, m in enumerate(mats)]
    dvs = [torch.from_numpy(v).cuda() for v in vs]
    b, report = tune(mats, dvs, reps=2)
    assert b.encoding in report and set(report) <= {"q32", "q64", "q128", "v0"}
    assert report[b.encoding] == min(report.values())
    b.launch()
    torch.cuda.synchronize()
    for m, v, o in zip(mats, vs, b.outputs):
        assert np.array_equal(o.cpu().numpy(), want(m, v))
