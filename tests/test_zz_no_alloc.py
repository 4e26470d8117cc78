"""The entry point that allocates nothing (matvans_v1_decompress_backward_noalloc),
called through the C ABI: its dx equals the workspace entry point's
(matvans_v1_decompress_backward) bit for bit, on two calls in a row, the second with a
smaller input (the static scratch keeps the first call's partial sums); a request larger
than the scratch is refused with MATVANS_E_CAPACITY before any launch."""
import ctypes
import sys

import numpy as np
import pytest

from matvans_amd import capi
from matvans_amd.capi import MatvansError, call
from matvans_amd.mvx import CompressedMatrixV1

from .helpers import gaussian_int8

pytestmark = pytest.mark.gpu


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def p(t):
    return ctypes.c_void_p(t.data_ptr())


def backward_both(cm, g, s_w, alpha):
    """(dx from the workspace entry point, dx from the no-alloc entry point)."""
    import torch
    n, rows = g.shape
    s = cm.dev_struct()
    need = int(capi.lib().matvans_backward_workspace_size(cm.rows, cm.cols, n))
    ws = torch.empty(need // 4 + 4, dtype=torch.float32, device="cuda")
    want = torch.full((n, cm.cols), float("nan"), device="cuda")
    got = torch.full((n, cm.cols), float("nan"), device="cuda")
    call("matvans_v1_decompress_backward", ctypes.byref(s), p(want), p(g), n, rows, p(s_w),
         p(alpha), p(ws), ws.numel() * 4, None)
    call("matvans_v1_decompress_backward_noalloc", ctypes.byref(s), p(got), p(g), n, rows,
         p(s_w), p(alpha), None)
    torch.cuda.synchronize()
    return want.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("streams", [32, 64])
def test_noalloc_backward_matches_workspace_entry_point(cuda, streams):
    import torch
    rng = np.random.default_rng(11)
    for rows, cols, n_vec in [(4096, 14336, 8), (300, 2048, 3)]:   # second call: smaller
        note(f"no-alloc backward q{streams} {rows}x{cols} n_vec={n_vec}")
        w, _ = gaussian_int8(rows, cols, 6.0, seed=rows + cols)
        cm = CompressedMatrixV1.encode(w, streams, 8, 0, coder=1).to_device()
        g = torch.from_numpy(rng.standard_normal((n_vec, rows)).astype(np.float32)).cuda()
        s_w = torch.from_numpy((rng.random(rows) * 0.02 + 0.001).astype(np.float32)).cuda()
        alpha = torch.from_numpy((rng.random(n_vec) + 0.5).astype(np.float32)).cuda()
        want, got = backward_both(cm, g, s_w, alpha)
        assert np.isfinite(want).all()
        assert want.tobytes() == got.tobytes(), (rows, cols, n_vec)


def test_noalloc_backward_refuses_oversized_request(cuda):
    import torch
    note("no-alloc backward capacity check")
    w, _ = gaussian_int8(64, 2048, 6.0, seed=3)
    cm = CompressedMatrixV1.encode(w, 32, 8, 0, coder=1).to_device()
    s = cm.dev_struct()
    s.rows = 16384                     # 8 x 256 row blocks x 2048 cols x 4 B = 16 MiB: fits
    s.cols = 8192                      # 8 x 256 x 8192 x 4 B = 64 MiB: too large
    dx = torch.full((8, 8192), float("nan"), device="cuda")
    g = torch.zeros((8, 16384), device="cuda")
    s_w = torch.ones(16384, device="cuda")
    alpha = torch.ones(8, device="cuda")
    with pytest.raises(MatvansError) as e:
        call("matvans_v1_decompress_backward_noalloc", ctypes.byref(s), p(dx), p(g), 8, 16384,
             p(s_w), p(alpha), None)
    assert e.value.status == 9
    assert bool(torch.isnan(dx).all())
