"""The Python bindings check every tensor they hand to the library (device, dtype,
shape, stride, length) and raise ValueError before any library call.

Host tensors are rejected as such.  The other checks run here on CPU tensors through the
device-test seam ``_check.is_device_tensor`` (patched to accept them) with the library's
``call`` patched to fail, so a rejection is proven to happen before anything reaches a
kernel."""
import numpy as np
import pytest
import torch

import matvans_amd as mv
from matvans_amd import _check, buffered, mvx


@pytest.fixture(scope="module")
def w():
    rng = np.random.default_rng(5)
    return np.clip(np.rint(rng.normal(0, 6, size=(32, 512))), -127, 127).astype(np.int8)


@pytest.fixture
def fake_device(monkeypatch):
    """CPU tensors count as device tensors; any library compute call fails the test."""
    monkeypatch.setattr(_check, "is_device_tensor", lambda t: True)

    def no_call(name, *a):
        raise AssertionError(f"{name} reached the library")
    for mod in (mv, mvx, buffered):
        monkeypatch.setattr(mod, "call", no_call)


@pytest.fixture(scope="module")
def m0(w):
    return mv.CompressedMatrix.encode(w).to_device("cpu")


@pytest.fixture(scope="module")
def mq(w):
    return mvx.CompressedMatrixV1.encode(w, 32, 8, 0, coder=1).to_device("cpu")


# ----------------------------------------------------------------- host tensors
def test_host_tensors_rejected(w, m0, mq):
    x = torch.zeros(512, dtype=torch.int8)
    for fn in (lambda: m0.multiply(x), lambda: mq.multiply(x),
               lambda: mv.int8_matvec(torch.from_numpy(w), x),
               lambda: mv.f32_gemv(torch.zeros(4, 8), torch.zeros(8)),
               lambda: mv.absmax_requant(torch.zeros(4, dtype=torch.int32)),
               lambda: mv.f32_normalize_round(torch.zeros(4)),
               lambda: mv.scale_output(torch.zeros(4, dtype=torch.int32), torch.zeros(4),
                                       torch.zeros(1)),
               lambda: mv.quantize_vectors(torch.zeros(8)),
               lambda: m0.backward(torch.zeros(32), torch.ones(32)),
               lambda: mq.backward(torch.zeros(32), torch.ones(32))):
        with pytest.raises(ValueError, match="must be a device tensor"):
            fn()


def test_check_rejects_non_tensor():
    with pytest.raises(ValueError, match="expected a torch tensor"):
        _check.tensor("x", np.zeros(4, np.int8), dtype=torch.int8)


# ----------------------------------------------------------------- forward
@pytest.mark.parametrize("which", ["v0", "quad"])
def test_multiply_rejections(fake_device, m0, mq, which):
    m = m0 if which == "v0" else mq
    with pytest.raises(ValueError, match="dtype"):
        m.multiply(torch.zeros(512, dtype=torch.int16))
    with pytest.raises(ValueError, match="expected \\[512\\]"):
        m.multiply(torch.zeros(256, dtype=torch.int8))
    with pytest.raises(ValueError, match="vectors, expected 1..8"):
        m.multiply(torch.zeros((9, 512), dtype=torch.int8))
    with pytest.raises(ValueError, match="contiguous"):
        m.multiply(torch.zeros(1024, dtype=torch.int8)[::2])
    with pytest.raises(ValueError, match="contiguous"):
        m.multiply(torch.zeros((512, 2), dtype=torch.int8).t())
    with pytest.raises(ValueError, match="out: shape"):
        m.multiply(torch.zeros(512, dtype=torch.int8), out=torch.zeros(16, dtype=torch.int32))
    with pytest.raises(ValueError, match="out: dtype"):
        m.multiply(torch.zeros(512, dtype=torch.int8), out=torch.zeros(32, dtype=torch.float32))


# ----------------------------------------------------------------- backward
@pytest.mark.parametrize("which", ["v0", "quad"])
def test_backward_rejections(fake_device, m0, mq, which):
    m = m0 if which == "v0" else mq
    g = torch.zeros((2, 32))
    with pytest.raises(ValueError, match="g: dtype"):
        m.backward(g.double(), torch.ones(32))
    with pytest.raises(ValueError, match="s_w: shape"):
        m.backward(g, torch.ones(31))
    with pytest.raises(ValueError, match="s_w: dtype"):
        m.backward(g, torch.ones(32, dtype=torch.float16))
    with pytest.raises(ValueError, match="alpha: shape"):
        m.backward(g, torch.ones(32), alpha=torch.ones(1))
    with pytest.raises(ValueError, match="out: shape"):
        m.backward(g, torch.ones(32), out=torch.zeros((2, 500)))
    with pytest.raises(ValueError, match="workspace: .*contiguous"):
        m.backward(g, torch.ones(32), workspace=torch.zeros((64, 64))[:, ::2])


# ----------------------------------------------------------------- dense ops
def test_dense_rejections(fake_device):
    mat = torch.zeros((8, 64), dtype=torch.int8)
    with pytest.raises(ValueError, match="vecs: shape"):
        mv.int8_matvec(mat, torch.zeros(32, dtype=torch.int8))
    with pytest.raises(ValueError, match="mat: .*contiguous"):
        mv.int8_matvec(torch.zeros((64, 8), dtype=torch.int8).t(), torch.zeros(64, dtype=torch.int8))
    with pytest.raises(ValueError, match="out: shape"):
        mv.int8_matvec(mat, torch.zeros(64, dtype=torch.int8), out=torch.zeros(4, dtype=torch.int32))
    with pytest.raises(ValueError, match="x: shape"):
        mv.f32_gemv(torch.zeros((4, 8)), torch.zeros(7))
    with pytest.raises(ValueError, match="out8: 2 elements"):
        mv.absmax_requant(torch.zeros(4, dtype=torch.int32), out8=torch.zeros(2, dtype=torch.int8))
    with pytest.raises(ValueError, match="delta: dtype"):
        mv.absmax_requant(torch.zeros(4, dtype=torch.int32), delta=torch.zeros(1, dtype=torch.int32))
    with pytest.raises(ValueError, match="y: dtype"):
        mv.f32_normalize_round(torch.zeros(4, dtype=torch.float64))
    with pytest.raises(ValueError, match="s_w: shape"):
        mv.scale_output(torch.zeros((2, 4), dtype=torch.int32), torch.zeros(3), torch.zeros(2))
    with pytest.raises(ValueError, match="s_x: shape"):
        mv.scale_output(torch.zeros((2, 4), dtype=torch.int32), torch.zeros(4), torch.zeros(1))
    with pytest.raises(ValueError, match="x: dtype"):
        mv.quantize_vectors(torch.zeros(8, dtype=torch.float64))


def test_reference_signature_rejections(fake_device, m0):
    args = dict(dst=torch.zeros(32, dtype=torch.int32), vec=torch.zeros(512, dtype=torch.int8),
                rows=32, cols=512, grid_spacing=1.0, cursors=m0.d_cursors, min_value=m0.min_value,
                G=m0.G, cdf_data=m0.d_cdf_data, ppf_data=m0.d_ppf_data, payload=m0.d_payload,
                payload_size=m0.payload_size)
    bad = dict(args, dst=torch.zeros(16, dtype=torch.int32))
    with pytest.raises(ValueError, match="dst: 16 elements"):
        mv.decmpressAndMultiply(**bad)
    bad = dict(args, vec=torch.zeros(512, dtype=torch.uint8))
    with pytest.raises(ValueError, match="vec: dtype"):
        mv.decmpressAndMultiply(**bad)
    bad = dict(args, payload_size=m0.payload_size + 10 ** 6)
    with pytest.raises(ValueError, match="payload: "):
        mv.decmpressAndMultiply(**bad)


def test_composite_rejections(fake_device, m0):
    with pytest.raises(ValueError, match="d_result8: dtype"):
        m0.decompress_and_mult(torch.zeros(32, dtype=torch.int32), torch.zeros(32, dtype=torch.int32),
                               torch.zeros(512, dtype=torch.int8), torch.zeros(1))
    with pytest.raises(ValueError, match="vector: 256 elements"):
        m0.decompress_and_mult(torch.zeros(32, dtype=torch.int8), torch.zeros(32, dtype=torch.int32),
                               torch.zeros(256, dtype=torch.int8), torch.zeros(1))


def test_batch_rejections(fake_device, m0, mq):
    from matvans_amd.batch import Batch, BatchBackward
    with pytest.raises(ValueError, match="entry 0 vecs: shape"):
        Batch([(m0, torch.zeros(100, dtype=torch.int8))])
    with pytest.raises(ValueError, match="entry 1 vecs: dtype"):
        Batch([(m0, torch.zeros(512, dtype=torch.int8)), (mq, torch.zeros(512, dtype=torch.int32))])
    with pytest.raises(ValueError, match="entry 0 s_w: shape"):
        BatchBackward([(mq, torch.zeros(32), torch.ones(16), None)])
    with pytest.raises(ValueError, match="entry 0 g: shape"):
        BatchBackward([(mq, torch.zeros(31), torch.ones(32), None)])
