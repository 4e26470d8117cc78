"""Persistent batched decode-matvec over the C ABI (matvans_batch_*): every matrix of a
batch -- .bin v0, MVX1 quad coder, family-B rows / stripes, uneven shapes, 1..8 vectors
each -- in one kernel launch fed by a device work queue, with an autotuner for the tile
size, tile order and persistent grid.  No reference counterpart (the reference launches
one kernel per matrix, matvansWarpPerRow.cu:156-164)."""
from __future__ import annotations

import ctypes

from . import _check
from .buffered import BufferedMatrices
from .capi import BatchBwdItem, BatchConfig, BatchItem, call, lib
from .mvx import CompressedMatrixV1

__all__ = ["Batch", "BatchBackward", "tune"]

V0, QUAD, B_ROWS, B_STRIPES, QUAD_LANES = 0, 1, 2, 3, 4


def _quad_format(m):
    """Batch format of an MVX1 matrix: the quad coder with one table per matrix (coder 1,
    or coder 2 with 32 or 64 streams), else None."""
    if not isinstance(m, CompressedMatrixV1) or m.table_mode != 0:
        return None
    if m.coder == 1:
        return QUAD
    if m.coder == 2 and m.streams in (32, 64):
        return QUAD_LANES
    return None


def _p(t):
    return t.data_ptr() if t is not None else None


class Batch:
    """entries: list of (matrix, vecs), (matrix, vecs, scale) or (matrix, vecs, scale,
    int32) where matrix is a CompressedMatrix (.bin v0), a CompressedMatrixV1 with
    coder=1, one table per matrix and 32, 64 or 128 streams (coder=2: 32 or 64 streams),
    or a (BufferedMatrices, k)
    pair (a ValueError names any other entry); vecs is an int8 device
    tensor [n_vec, cols] (or [cols]).  All matrices must be on the device.
    ``outputs[i]`` is the int32 [n_vec, rows] result of entry i (None when int32 is
    False); ``out8[i]`` the next layer's int8 input clamp(rint(float(acc) * scale)),
    computed in the same launch, for entries with a scale (family-B entries: always, with
    the file's scale unless one is given), else None."""

    def __init__(self, entries):
        import torch
        from . import CompressedMatrix
        self._keep = []
        self.outputs, self.out8 = [], []
        arr = (BatchItem * len(entries))()
        for i, entry in enumerate(entries):
            m, v, *rest = entry
            scale = rest[0] if rest else None
            want32 = bool(rest[1]) if len(rest) > 1 else True
            if not want32 and scale is None:
                raise ValueError("batch entry %d: no output (int32 off and no scale)" % i)
            if isinstance(m, tuple):
                bm, k = m
                if not 0 <= k < bm.num:
                    raise ValueError(f"batch entry {i}: matrix index {k} outside 0..{bm.num - 1}")
                if bm.d is None:
                    raise ValueError(f"batch entry {i}: matrices are not on the device")
                mdev, mcols = bm.d["compressed"].device, bm.shapes[k][1]
            else:
                if not (isinstance(m, CompressedMatrix) or _quad_format(m) is not None):
                    raise ValueError("batch entries: .bin v0, MVX1 quad (one table per "
                                     "matrix) or family-B matrices")
                if getattr(m, "d_payload", None) is None:
                    raise ValueError(f"batch entry {i}: matrix is not on the device")
                mdev, mcols = m.d_payload.device, m.cols
            v = _check.vectors(f"entry {i} vecs", v, cols=mcols, device=mdev, dtype=torch.int8)
            self._keep.append(v)
            it = arr[i]
            it.n_vec, it.vec_stride, it.vecs = v.shape[0], v.stride(0), v.data_ptr()
            if isinstance(m, tuple):
                bm, k = m
                rows, cols = bm.shapes[k]
                acc = torch.empty((v.shape[0], rows), dtype=torch.int32, device=v.device)
                o8 = torch.empty((v.shape[0], rows), dtype=torch.int8, device=v.device)
                it.format = B_ROWS if bm.layout == "rows" else B_STRIPES
                it.compressed = _p(bm.d["compressed"])
                it.compressed_words = bm.compressed.size
                it.warp_offsets = _p(bm.d["offsets"][k])
                it.end_offset = bm.end_offset(k)
                it.b_ppf = bm.d["ppf"].data_ptr() + 128 * k
                it.b_cum = bm.d["cum"].data_ptr() + 128 * k
                it.b_prob = bm.d["prob"].data_ptr() + 128 * k
                it.scale = float(bm.scales[k] if scale is None else scale)
                if not want32:
                    acc = None
                it.acc32, it.out8 = _p(acc), o8.data_ptr()
                it.dst = _p(acc)
            else:
                rows, cols = m.rows, m.cols
                acc = (torch.empty((v.shape[0], rows), dtype=torch.int32, device=v.device)
                       if want32 else None)
                o8 = None
                if scale is not None:
                    o8 = torch.empty((v.shape[0], rows), dtype=torch.int8, device=v.device)
                    it.out8, it.scale = o8.data_ptr(), float(scale)
                it.dst, it.cursors = _p(acc), _p(m.d_cursors)
                it.payload, it.payload_size = _p(m.d_payload), m.payload_size
                if isinstance(m, CompressedMatrix):
                    it.format = V0
                    it.min_value, it.G = m.min_value, m.G
                    it.cdf, it.ppf = _p(m.d_cdf_data), _p(m.d_ppf_data)
                elif _quad_format(m) is not None:
                    it.format, it.streams, it.lut = _quad_format(m), m.streams, _p(m.d_lut)
                    if m.d_out is not None:
                        it.out_start, it.out_rec = _p(m.d_out[0]), _p(m.d_out[1])
                else:
                    raise ValueError("batch entries: .bin v0, MVX1 quad (one table per "
                                     "matrix) or family-B matrices")
            it.rows, it.cols = rows, cols
            self.outputs.append(acc)
            self.out8.append(o8)
            self._keep.append(m)
        self._items = arr
        h = ctypes.c_void_p()
        call("matvans_batch_create", arr, len(entries), ctypes.byref(h))
        self._h = h

    def close(self):
        """Release the plan's device memory (also done when the object is collected)."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().matvans_batch_destroy(h)
        self._h = None

    def __del__(self):
        # at interpreter exit the module globals may already be gone; the process exit
        # releases the device memory then
        try:
            self.close()
        except (TypeError, AttributeError):
            pass

    def configure(self, units_per_tile=1, order=0, blocks_per_cu=8, schedule=0):
        """schedule 0: blocks take tiles from the device work queue; 1: block b takes
        tiles b, b + grid, b + 2 grid, ... (no atomics)."""
        cfg = BatchConfig(units_per_tile, order, blocks_per_cu, schedule)
        call("matvans_batch_configure", self._h, ctypes.byref(cfg))

    @property
    def config(self):
        cfg = BatchConfig()
        call("matvans_batch_get_config", self._h, ctypes.byref(cfg))
        return (cfg.units_per_tile, cfg.order, cfg.blocks_per_cu, cfg.schedule)

    def launch(self, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        call("matvans_batch_launch", self._h, ctypes.c_void_p(s.cuda_stream))

    def autotune(self, reps=5, stream=None):
        """Returns ((units_per_tile, order, blocks_per_cu, schedule), ms per launch) of the
        best."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        cfg = BatchConfig()
        ms = ctypes.c_float()
        call("matvans_batch_autotune", self._h, ctypes.c_void_p(s.cuda_stream), reps,
             ctypes.byref(cfg), ctypes.byref(ms))
        return (cfg.units_per_tile, cfg.order, cfg.blocks_per_cu, cfg.schedule), ms.value


class BatchBackward(Batch):
    """The backward dx = alpha * W^T (s_w * g) of every matrix of a batch in two launches
    (persistent partial sums from the work queue, then one fixed-order reduce); bitwise
    equal to each matrix's own ``backward``.  entries: list of (matrix, g, s_w, alpha)
    with matrix a CompressedMatrix (.bin v0) or a quad CompressedMatrixV1 (32 or 64
    streams, one table per matrix) on the device, g fp32 [n_vec, rows] (or [rows]), s_w
    fp32 [rows], alpha
    fp32 [n_vec] or None (ones).  ``outputs[i]`` is dx of entry i, fp32 [n_vec, cols]."""

    def __init__(self, entries):  # noqa: D107 (the base initialiser is not used)
        import torch
        from . import CompressedMatrix
        self._keep = []
        self.outputs, self.out8 = [], []
        arr = (BatchBwdItem * len(entries))()
        for i, (m, g, s_w, alpha) in enumerate(entries):
            if not (isinstance(m, CompressedMatrix) or _quad_format(m) is not None):
                raise ValueError("backward batch entries: .bin v0 or MVX1 quad matrices "
                                 "(one table per matrix)")
            if getattr(m, "d_payload", None) is None:
                raise ValueError(f"backward batch entry {i}: matrix is not on the device")
            dev = m.d_payload.device
            gg = _check.vectors(f"entry {i} g", g, cols=m.rows, device=dev, dtype=torch.float32)
            n = gg.shape[0]
            _check.tensor(f"entry {i} s_w", s_w, dtype=torch.float32, device=dev, shape=(m.rows,))
            if alpha is None:
                alpha = torch.ones(n, dtype=torch.float32, device=dev)
            _check.tensor(f"entry {i} alpha", alpha, dtype=torch.float32, device=dev, shape=(n,))
            dx = torch.empty((n, m.cols), dtype=torch.float32, device=dev)
            it = arr[i]
            it.rows, it.cols, it.n_vec, it.g_stride = m.rows, m.cols, n, gg.stride(0)
            it.g, it.s_w, it.alpha, it.dx = gg.data_ptr(), s_w.data_ptr(), alpha.data_ptr(), dx.data_ptr()
            it.cursors, it.payload, it.payload_size = _p(m.d_cursors), _p(m.d_payload), m.payload_size
            if isinstance(m, CompressedMatrix):
                it.format = V0
                it.min_value, it.G = m.min_value, m.G
                it.cdf, it.ppf = _p(m.d_cdf_data), _p(m.d_ppf_data)
            elif _quad_format(m) is not None:
                it.format, it.streams, it.lut = _quad_format(m), m.streams, _p(m.d_lut)
                if m.d_out is not None:
                    it.out_col_start, it.out_col_rec = _p(m.d_out[2]), _p(m.d_out[3])
            else:
                raise ValueError("backward batch entries: .bin v0 or MVX1 quad matrices "
                                 "(one table per matrix)")
            self._keep += [m, gg, s_w, alpha]
            self.outputs.append(dx)
            self.out8.append(None)
        self._items = arr
        h = ctypes.c_void_p()
        call("matvans_batch_create_backward", arr, len(entries), ctypes.byref(h))
        self._h = h


def tune(matrices, vecs, encodings=("q32", "q64", "q128", "v0"), reps=5, stream=None):
    """Pick the encoding (streams per row, and with it rows per wavefront: two rows per
    wave at 32 streams, one at 64, half a row at 128) and the plan configuration (tile size, tile order,
    persistent grid, schedule) for a batch of raw int8 matrices.  matrices: int8 numpy
    arrays [rows, cols]; vecs: int8 device tensors [n_vec, cols], one per matrix.  Each
    candidate encoding is encoded on the host, uploaded, and its plan autotuned; the
    fastest plan is returned with a report {encoding: ms per launch}.  The other
    candidates' device memory is released."""
    from . import CompressedMatrix
    import torch
    report, best = {}, None
    dev = vecs[0].device
    for enc in encodings:
        if enc == "v0":
            mats = [CompressedMatrix.encode(m).to_device(dev) for m in matrices]
        elif enc in ("q32", "q64", "q128", "l32", "l64"):
            S = int(enc[1:])
            if any(m.shape[1] % (16 * S) for m in matrices):
                continue
            coder = 2 if enc[0] == "l" else 1
            mats = [CompressedMatrixV1.encode(m, S, 8, 0, coder=coder).to_device(dev)
                    for m in matrices]
        else:
            raise ValueError(f"unknown encoding {enc}")
        b = Batch(list(zip(mats, vecs)))
        _, ms = b.autotune(reps=reps, stream=stream)
        report[enc] = ms
        if best is None or ms < best[1]:
            if best is not None:
                best[0].close()
            best = (b, ms, enc)
        else:
            b.close()
        torch.cuda.synchronize()
    if best is None:
        raise ValueError("no candidate encoding fits the matrices' shapes")
    best[0].encoding = best[2]
    return best[0], report
