"""Argument validation of the batch plan constructors (matvans_batch_create,
matvans_batch_create_backward): every rejected item is refused before any device call,
so these run without a GPU."""
import ctypes

import pytest

from matvans_amd.capi import BatchBwdItem, BatchItem, MatvansError, call

V0, QUAD, B_ROWS, B_STRIPES, LANES = 0, 1, 2, 3, 4


def fwd_item(**kw):
    it = BatchItem()
    it.format, it.rows, it.cols, it.n_vec, it.vec_stride = QUAD, 64, 2048, 1, 2048
    it.vecs, it.dst, it.cursors, it.payload, it.lut, it.streams = 16, 16, 16, 16, 16, 32
    for k, v in kw.items():
        setattr(it, k, v)
    return it


def bwd_item(**kw):
    it = BatchBwdItem()
    it.format, it.rows, it.cols, it.n_vec, it.g_stride = QUAD, 64, 2048, 1, 64
    it.g, it.s_w, it.alpha, it.dx = 16, 16, 16, 16
    it.cursors, it.payload, it.lut, it.streams = 16, 16, 16, 32
    for k, v in kw.items():
        setattr(it, k, v)
    return it


def create(fn, items):
    arr = (type(items[0]) * len(items))(*items)
    h = ctypes.c_void_p()
    with pytest.raises(MatvansError) as e:
        call(fn, arr, len(items), ctypes.byref(h))
    assert not h.value
    return e.value.status


@pytest.mark.parametrize("kw,status", [
    ({"vecs": 0}, 1),                        # null vectors
    ({"dst": 0}, 1),                         # quad: neither int32 nor int8 output
    ({"format": V0, "dst": 0}, 1),           # .bin v0: the same
    ({"n_vec": 9}, 1),                       # more than 8 vectors
    ({"n_vec": 0}, 1),
    ({"rows": 0}, 2),                        # empty matrix
    ({"streams": 96}, 8),                    # quad items: 32, 64 or 128 streams
    ({"cols": 1000}, 2),                     # cols not a multiple of 16 * streams
    ({"payload": 18}, 3),                    # payload not 16-byte aligned
    ({"vecs": 20}, 3),                       # vectors not 16-byte aligned
    ({"format": V0, "cols": 2000}, 2),       # .bin v0: cols % 512
    ({"format": V0, "cdf": 0}, 1),           # .bin v0 tables missing
    ({"format": B_STRIPES, "rows": 40}, 2),  # stripes: rows % 32
    ({"format": 7}, 1),                      # unknown format
    ({"format": LANES, "streams": 128}, 8),  # coder 2: 32 or 64 streams
    ({"format": LANES, "payload_size": 64 * 34 - 1}, 2),   # payload shorter than row tables
])
def test_forward_plan_rejects(kw, status):
    assert create("matvans_batch_create", [fwd_item(), fwd_item(**kw)]) == status


@pytest.mark.parametrize("kw,status", [
    ({"g": 0}, 1),
    ({"n_vec": 9}, 1),
    ({"n_vec": 2, "g_stride": 10}, 1),       # g_stride < rows
    ({"rows": 0}, 2),
    ({"streams": 128}, 8),
    ({"cols": 1000}, 2),
    ({"payload": 24}, 3),
    ({"format": V0, "cdf": 0}, 1),
    ({"format": B_ROWS}, 8),                 # family-B items have no backward
    ({"format": LANES, "payload_size": 64 * 34 - 1}, 2),   # payload shorter than row tables
])
def test_backward_plan_rejects(kw, status):
    assert create("matvans_batch_create_backward", [bwd_item(), bwd_item(**kw)]) == status


def test_empty_and_null_plans():
    h = ctypes.c_void_p()
    with pytest.raises(MatvansError) as e:
        call("matvans_batch_create", (BatchItem * 1)(), 0, ctypes.byref(h))
    assert e.value.status == 1
    with pytest.raises(MatvansError) as e:
        call("matvans_batch_launch", None, None)
    assert e.value.status == 1


def test_per_row_table_matrices_are_refused_by_name():
    """A quad matrix with one table per row has no matrix-wide decode table: both batch
    constructors name the accepted encodings instead of failing on a null table."""
    import numpy as np
    import torch
    from matvans_amd.batch import Batch, BatchBackward
    from matvans_amd.mvx import CompressedMatrixV1
    m = np.random.default_rng(1).integers(-20, 21, (64, 1024)).astype(np.int8)
    rowcdf = CompressedMatrixV1.encode(m, 64, 8, 1, coder=1)
    single = CompressedMatrixV1.encode(m, 64, 8, 0, coder=0)
    for c in (rowcdf, single):
        with pytest.raises(ValueError, match="one table per matrix"):
            Batch([(c, torch.zeros(1024, dtype=torch.int8))])
        with pytest.raises(ValueError, match="one table per matrix"):
            BatchBackward([(c, torch.zeros(64), torch.ones(64), None)])
