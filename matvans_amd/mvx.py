"""MVX1 matrices: the format extensions of the WarpPerRow codec (64 streams per row,
12-bit probabilities, one table per row), over the C ABI of ``lib/libmatvans.so``.

  CompressedMatrixV1.encode(W, streams, precision, table_mode)   library encoder
  CompressedMatrixV1.to_device() / .multiply(vecs)               fused decode-matvec
  read_mvx1(path) / write_mvx1(path, ...)                        MVX1 container

``streams=32, precision=8, table_mode=0`` is the reference's layout and model
(create_matrices_WarpPerRow.py:131-272,353-430); the others are extensions.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check
from .capi import V1DevMatrix, V1Host, call, lib

__all__ = ["CompressedMatrixV1", "read_mvx1", "write_mvx1"]


def _np_ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class CompressedMatrixV1:
    """One MVX1 matrix.  Host arrays: cursors uint32[rows], cdf uint16[G+1] (table_mode 0)
    or freq uint8[rows, G] (table_mode 1), payload uint16[payload_size]."""

    def __init__(self, rows, cols, streams, precision, table_mode, min_value, G, cursors,
                 payload, cdf=None, freq=None, grid_spacing=1.0, coder=0, out_start=None,
                 out_rec=None):
        self.coder = int(coder)
        self.rows, self.cols = int(rows), int(cols)
        self.streams, self.precision, self.table_mode = int(streams), int(precision), int(table_mode)
        self.min_value, self.G = int(min_value), int(G)
        self.cursors = np.ascontiguousarray(cursors, np.uint32)
        self.payload = np.ascontiguousarray(payload, np.uint16)
        self.payload_size = len(self.payload)
        self.cdf = None if cdf is None else np.ascontiguousarray(cdf, np.uint16)
        self.freq = None if freq is None else np.ascontiguousarray(freq, np.uint8).reshape(
            self.rows, self.G)
        self.grid_spacing = float(grid_spacing)
        # outlier split: entries coded as 0, listed by row (col | (uint8)value << 24)
        self.out_start = None if out_start is None else np.ascontiguousarray(out_start, np.uint32)
        self.out_rec = None if out_rec is None else np.ascontiguousarray(out_rec, np.uint32)
        self.n_outliers = 0 if self.out_rec is None else int(self.out_rec.size)
        self.d_cursors = self.d_lut = self.d_freq = self.d_payload = None
        self.d_out = None

    @property
    def stored_bytes(self):
        """Payload + cursors + tables (+ outlier offsets and records), as stored."""
        tab = 2 * (self.G + 1) if self.table_mode == 0 else self.rows * self.G
        out = 4 * (self.rows + 1) + 4 * self.n_outliers if self.n_outliers else 0
        return 2 * self.payload_size + 4 * self.rows + tab + out

    @staticmethod
    def outlier_band(matrix):
        """The band T of the outlier split the library picks for this matrix (127: none)."""
        m = np.ascontiguousarray(np.asarray(matrix), dtype=np.int8)
        t = ctypes.c_int32()
        call("matvans_v1_outlier_band", _np_ptr(m), m.size, ctypes.byref(t))
        return t.value

    @staticmethod
    def split_outliers(matrix, band):
        """(coded matrix with |w| > band set to 0, out_start uint32[rows + 1],
        out_rec uint32[n] = col | (uint8)w << 24 in row-major order)."""
        m = np.ascontiguousarray(np.asarray(matrix), dtype=np.int8)
        rows, cols = m.shape
        cap = int(np.count_nonzero(np.abs(m.astype(np.int16)) > band))
        coded = np.empty_like(m)
        start = np.empty(rows + 1, np.uint32)
        rec = np.empty(max(1, cap), np.uint32)
        n = ctypes.c_uint64()
        call("matvans_v1_split_outliers", _np_ptr(m), rows, cols, int(band), _np_ptr(coded),
             _np_ptr(start), _np_ptr(rec), cap, ctypes.byref(n))
        return coded, start, rec[:n.value].copy()

    @classmethod
    def encode(cls, matrix, streams=64, precision=8, table_mode=0, n_threads=0, coder=0,
               outliers=None):
        """coder 0: one symbol per step; 1: quad coder (P = 8); 2: quad coder with the
        streams back to back (P = 8, one table; no per-stream padding).
        outliers (quad coder, one table): None or False, no split; "auto", the band the
        library picks; an int, that band."""
        m = np.ascontiguousarray(np.asarray(matrix), dtype=np.int8)
        rows, cols = m.shape
        out_start = out_rec = None
        if outliers is not None and outliers is not False:
            if coder not in (1, 2) or table_mode != 0 or precision != 8:
                raise ValueError("the outlier split needs the quad coder with one 8-bit table")
            band = cls.outlier_band(m) if outliers == "auto" else int(outliers)
            if band < 127:
                m, out_start, out_rec = cls.split_outliers(m, band)
                if out_rec.size == 0:
                    out_start = out_rec = None
        mn, G = ctypes.c_int32(), ctypes.c_uint32()
        cdf = np.zeros(257, np.uint16)
        freq = np.zeros(rows * 256 if table_mode == 1 else 1, np.uint8)
        call("matvans_v1_model", _np_ptr(m), rows, cols, precision, table_mode, ctypes.byref(mn),
             ctypes.byref(G), _np_ptr(cdf), _np_ptr(freq))
        g = G.value
        cap = int(lib().matvans_v1_encode_bound(rows, cols, streams, precision))
        cur = np.empty(rows, np.uint32)
        pay = np.empty(cap, np.uint16)
        size = ctypes.c_uint64()
        call("matvans_v1_encode", _np_ptr(m), rows, cols, streams, precision, table_mode,
             coder, mn.value, g, _np_ptr(cdf), _np_ptr(freq), int(n_threads), _np_ptr(cur),
             _np_ptr(pay), cap, ctypes.byref(size))
        return cls(rows, cols, streams, precision, table_mode, mn.value, g, cur,
                   pay[:size.value].copy(),
                   cdf=cdf[:g + 1].copy() if table_mode == 0 else None,
                   freq=freq[:rows * g].copy() if table_mode == 1 else None, coder=coder,
                   out_start=out_start, out_rec=out_rec)

    def outliers_by_column(self):
        """(out_col_start uint32[cols + 1], out_col_rec uint32[n] = row | (uint8)w << 24,
        rows ascending within a column): the records as the backward reads them."""
        rows_of = np.repeat(np.arange(self.rows, dtype=np.uint32), np.diff(self.out_start))
        cols = self.out_rec & 0xFFFFFF
        order = np.lexsort((rows_of, cols))
        col_start = np.zeros(self.cols + 1, np.uint32)
        np.cumsum(np.bincount(cols, minlength=self.cols), out=col_start[1:])
        col_rec = rows_of[order] | (self.out_rec[order] & np.uint32(0xFF000000))
        return col_start, col_rec.astype(np.uint32)

    def lut(self):
        """Decode table of a one-table matrix (2 * 2^P uint32)."""
        out = np.zeros(2 << self.precision, np.uint32)
        call("matvans_v1_lut", _np_ptr(self.cdf), self.G, self.min_value, self.precision,
             _np_ptr(out))
        return out

    def to_device(self, device="cuda"):
        import torch
        dev = torch.device(device)
        self.d_cursors = torch.from_numpy(self.cursors.view(np.int32)).to(dev)
        self.d_payload = torch.from_numpy(self.payload.view(np.int16)).to(dev)
        if self.table_mode == 0:
            self.d_lut = torch.from_numpy(self.lut().view(np.int32)).to(dev)
        else:
            self.d_freq = torch.from_numpy(self.freq.reshape(-1)).to(dev)
        if self.n_outliers:
            cs, cr = self.outliers_by_column()
            t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to(dev)
            self.d_out = (t(self.out_start), t(self.out_rec), t(cs), t(cr))
        return self

    def dev_struct(self) -> V1DevMatrix:
        if self.d_payload is None:
            raise RuntimeError("matrix is not on the device; call to_device() first")
        o = [t.data_ptr() for t in self.d_out] if self.d_out is not None else [0, 0, 0, 0]
        return V1DevMatrix(self.rows, self.cols, self.streams, self.precision, self.table_mode,
                           self.coder, self.min_value, self.G, self.payload_size, self.d_cursors.data_ptr(),
                           self.d_lut.data_ptr() if self.d_lut is not None else 0,
                           self.d_freq.data_ptr() if self.d_freq is not None else 0,
                           self.d_payload.data_ptr(), self.n_outliers, *o)

    def multiply(self, vec, out=None, stream=None):
        """int32 W @ vec for one vector [cols] or a batch [n_vec, cols] (n_vec <= 8)."""
        import torch
        s = self.dev_struct()
        dev = self.d_payload.device
        squeeze = vec.dim() == 1
        v = _check.vectors("vec", vec, cols=self.cols, device=dev, dtype=torch.int8)
        n = v.shape[0]
        if out is None:
            out = torch.empty((n, self.rows), dtype=torch.int32, device=dev)
        _check.output("out", out, shape=(n, self.rows), dtype=torch.int32, device=dev)
        call("matvans_v1_decompress_and_multiply", ctypes.byref(s), _ptr(out), _ptr(v), n,
             v.stride(0), _stream(stream))
        return out[0] if squeeze else out

    def backward(self, g, s_w, alpha=None, out=None, workspace=None, stream=None):
        """dx[v] = alpha[v] * W^T (s_w * g[v]) re-decoded from the quad-coder streams
        (coder 1 or 2, 32, 64 or 128 streams), g [rows] or [n_vec, rows] fp32 (n_vec <= 8);
        bitwise reproducible."""
        import torch
        s = self.dev_struct()
        dev = self.d_payload.device
        squeeze = g.dim() == 1
        gg = _check.vectors("g", g, cols=self.rows, device=dev, dtype=torch.float32)
        n = gg.shape[0]
        _check.tensor("s_w", s_w, dtype=torch.float32, device=dev, shape=(self.rows,))
        if alpha is None:
            alpha = torch.ones(n, dtype=torch.float32, device=dev)
        _check.tensor("alpha", alpha, dtype=torch.float32, device=dev, shape=(n,))
        need = int(lib().matvans_backward_workspace_size(self.rows, self.cols, n))
        if workspace is not None:
            _check.tensor("workspace", workspace, dtype=None, device=dev,
                          dtypes=(torch.float32, torch.int32, torch.uint8, torch.int8))
            if not workspace.is_contiguous():
                raise ValueError("workspace: must be contiguous")
        if workspace is None or workspace.numel() * workspace.element_size() < need:
            workspace = torch.empty(max(1, need // 4), dtype=torch.float32, device=dev)
        if out is None:
            out = torch.empty((n, self.cols), dtype=torch.float32, device=dev)
        _check.output("out", out, shape=(n, self.cols), dtype=torch.float32, device=dev)
        call("matvans_v1_decompress_backward", ctypes.byref(s), _ptr(out), _ptr(gg), n,
             gg.stride(0), _ptr(s_w), _ptr(alpha), _ptr(workspace),
             workspace.numel() * workspace.element_size(), _stream(stream))
        return out[0] if squeeze else out

    def host_struct(self, keep) -> V1Host:
        keep += [self.cursors, self.payload, self.cdf, self.freq, self.out_start, self.out_rec]
        return V1Host(self.rows, self.cols, self.grid_spacing, self.streams, self.precision,
                      self.table_mode, self.coder, self.min_value, self.G, self.payload_size,
                      self.cursors.ctypes.data,
                      self.cdf.ctypes.data if self.cdf is not None else 0,
                      self.freq.ctypes.data if self.freq is not None else 0,
                      self.payload.ctypes.data, self.n_outliers,
                      self.out_start.ctypes.data if self.n_outliers else 0,
                      self.out_rec.ctypes.data if self.n_outliers else 0)


def read_mvx1(path):
    """Returns (header dict, vector int8, [CompressedMatrixV1])."""
    h = ctypes.c_void_p()
    call("matvans_v1_read", str(path).encode(), ctypes.byref(h))
    try:
        num, rh, lv = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        vp = ctypes.c_void_p()
        call("matvans_v1_header", h, ctypes.byref(num), ctypes.byref(rh), ctypes.byref(vp),
             ctypes.byref(lv))

        def arr(p, ct, n):
            if n == 0:
                return np.zeros(0, np.dtype(ct))
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), (n,)).copy()
        vec = arr(vp, ctypes.c_int8, lv.value)
        mats = []
        for k in range(num.value):
            m = V1Host()
            call("matvans_v1_matrix", h, k, ctypes.byref(m))
            mats.append(CompressedMatrixV1(
                m.rows, m.cols, m.streams, m.precision, m.table_mode, m.min_value, m.G,
                arr(m.cursors, ctypes.c_uint32, m.rows),
                arr(m.payload, ctypes.c_uint16, m.payload_size),
                cdf=arr(m.cdf, ctypes.c_uint16, m.G + 1) if m.table_mode == 0 else None,
                freq=arr(m.freq, ctypes.c_uint8, m.rows * m.G) if m.table_mode == 1 else None,
                grid_spacing=m.grid_spacing, coder=m.coder,
                out_start=arr(m.out_start, ctypes.c_uint32, m.rows + 1) if m.n_outliers else None,
                out_rec=arr(m.out_rec, ctypes.c_uint32, m.n_outliers) if m.n_outliers else None))
        return {"num_matrices": num.value, "result_hash": rh.value}, vec, mats
    finally:
        lib().matvans_v1_free(h)


def write_mvx1(path, matrices, vector=None, result_hash=0):
    vec = np.ascontiguousarray(np.zeros(0, np.int8) if vector is None else vector, np.int8)
    keep = []
    arr = (V1Host * max(1, len(matrices)))()
    for i, m in enumerate(matrices):
        arr[i] = m.host_struct(keep)
    call("matvans_v1_write", str(path).encode(), result_hash, _np_ptr(vec), len(vec),
         len(matrices), arr)
