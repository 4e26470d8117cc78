"""MI355X-native fused rANS-decode x int8 matvec (drop-in for MatVANS WarpPerRow).

Host-side mirror of the reference's operator interface over the C ABI of
``lib/libmatvans.so`` (``include/matvans.h``):

  CompressedMatrix            matvansWarpPerRow/matvansWarpPerRow.cu:75-176
    .deserialize_file(path)   the .bin reader (:100-145, :225-258)
    .decompressAndMult(...)   decode-matvec + absmax + requant (:150-176)
  decmpressAndMultiply(...)   the kernel entry (matvansWarpPerRow-kernel.cu:46-50)
  hash_int8_array(a)          FxHash (matvansWarpPerRow.cu:31-42)
  run_chain(...)              the chained benchmark loop of main (:280-367)

PyTorch provides device memory and streams only.  Every compute call goes through the
HIP library; a missing library raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi
from . import _check
from .capi import DevMatrix, HostMatrix, MatvansError, call, lib

__all__ = ["CompressedMatrix", "MatvansError", "decmpressAndMultiply", "decompress_and_multiply",
           "int8_matvec", "f32_gemv", "absmax_requant", "f32_normalize_round", "scale_output",
           "hash_int8_array",
           "build_model", "encode_rows", "read_bin", "write_bin", "run_chain", "lib",
           "quantize_vectors", "compressed_linear"]


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _np_ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# ----------------------------------------------------------------------- host codec
def hash_int8_array(arr) -> int:
    """FxHash of an int8 array (matvansWarpPerRow.cu:31-42)."""
    a = np.ascontiguousarray(np.asarray(arr), dtype=np.int8)
    return int(lib().matvans_fxhash(_np_ptr(a), a.size))


def build_model(data):
    """8-bit entropy model (create_matrices_WarpPerRow.py:131-208).
    Returns (min_value, G, cdf16 uint16[G+1], ppf uint8[256])."""
    d = np.ascontiguousarray(np.asarray(data), dtype=np.int8).ravel()
    mn = ctypes.c_int8()
    G = ctypes.c_uint32()
    cdf = np.zeros(257, np.uint16)
    ppf = np.zeros(256, np.uint8)
    call("matvans_build_model", _np_ptr(d), d.size, ctypes.byref(mn), ctypes.byref(G),
         _np_ptr(cdf), _np_ptr(ppf))
    return int(mn.value), int(G.value), cdf[:G.value + 1].copy(), ppf


def encode_rows(matrix, min_value, cdf16, n_threads=0):
    """32-stream row-interleaved encode (create_matrices_WarpPerRow.py:353-430).
    Returns (cursors uint32[rows], payload uint16[payload_size])."""
    m = np.ascontiguousarray(np.asarray(matrix), dtype=np.int8)
    rows, cols = m.shape
    c16 = np.ascontiguousarray(cdf16, np.uint16)
    cap = int(lib().matvans_encode_bound(rows, cols))
    cur = np.empty(rows, np.uint32)
    pay = np.empty(cap, np.uint16)
    size = ctypes.c_uint64()
    call("matvans_encode_rows", _np_ptr(m), rows, cols, int(min_value), len(c16) - 1,
         _np_ptr(c16), int(n_threads), _np_ptr(cur), _np_ptr(pay), cap, ctypes.byref(size))
    return cur, pay[:size.value].copy()


class CompressedMatrix:
    """One rANS-compressed int8 matrix (matvansWarpPerRow.cu:75-148).

    Host arrays: cursors (uint32[rows]), cdf_data (uint8[G+1], cdf & 0xFF), ppf_data
    (uint8[256]), payload (uint16[payload_size]).  ``to_device`` uploads them as torch
    tensors (``d_cursors``, ``d_cdf_data``, ``d_ppf_data``, ``d_payload``)."""

    # words of zero slack after the payload; reads never reach it, but a caller-built
    # payload may be sliced by row and this keeps 16-byte loads inside the allocation
    SLACK_WORDS = 512

    def __init__(self, rows, cols, grid_spacing, cursors, min_value, G, cdf_data, ppf_data,
                 payload, payload_size=None):
        self.rows, self.cols = int(rows), int(cols)
        self.grid_spacing = float(grid_spacing)
        self.cursors = np.ascontiguousarray(cursors, np.uint32)
        self.min_value, self.G = int(min_value), int(G)
        self.cdf_data = np.ascontiguousarray(cdf_data, np.uint8)
        self.ppf_data = np.ascontiguousarray(ppf_data, np.uint8)
        self.payload = np.ascontiguousarray(payload, np.uint16)
        self.payload_size = int(len(self.payload) if payload_size is None else payload_size)
        self.d_cursors = self.d_cdf_data = self.d_ppf_data = self.d_payload = None

    # -- construction
    @classmethod
    def encode(cls, matrix, model=None, n_threads=0):
        """Encode an int8 matrix (rows x cols, cols % 512 == 0) with the library encoder.
        ``model`` = (min_value, cdf16) reuses a given model; otherwise it is built."""
        m = np.ascontiguousarray(np.asarray(matrix), dtype=np.int8)
        if model is None:
            mn, G, cdf16, ppf = build_model(m)
        else:
            mn, cdf16 = model
            cdf16 = np.asarray(cdf16, np.uint16)
            G = len(cdf16) - 1
            ppf = _ppf_from_cdf(cdf16)
        cur, pay = encode_rows(m, mn, cdf16, n_threads)
        cdf8 = (cdf16.astype(np.int64) & 0xFF).astype(np.uint8)
        return cls(m.shape[0], m.shape[1], 1.0, cur, mn, G, cdf8, ppf, pay)

    @staticmethod
    def deserialize_file(path):
        """Read a .bin v0 file.  Returns (header dict, vector int8, [CompressedMatrix])."""
        return read_bin(path)

    # -- device
    def to_device(self, device="cuda"):
        torch = _torch()
        dev = torch.device(device)
        self.d_cursors = torch.from_numpy(self.cursors.view(np.int32)).to(dev)
        cdf = np.zeros(257, np.uint8)
        cdf[:len(self.cdf_data)] = self.cdf_data
        self.d_cdf_data = torch.from_numpy(cdf).to(dev)
        self.d_ppf_data = torch.from_numpy(self.ppf_data.copy()).to(dev)
        pay = np.zeros(self.payload_size + self.SLACK_WORDS, np.uint16)
        pay[:self.payload_size] = self.payload[:self.payload_size]
        self.d_payload = torch.from_numpy(pay.view(np.int16)).to(dev)
        return self

    def dev_struct(self) -> DevMatrix:
        if self.d_payload is None:
            raise RuntimeError("matrix is not on the device; call to_device() first")
        return DevMatrix(self.rows, self.cols, self.grid_spacing, self.min_value, self.G, 0,
                         self.payload_size, self.d_cursors.data_ptr(),
                         self.d_cdf_data.data_ptr(), self.d_ppf_data.data_ptr(),
                         self.d_payload.data_ptr())

    def multiply(self, vec, out=None, stream=None):
        """int32 W @ vec for one vector [cols] or a batch [n_vec, cols] (n_vec <= 8)."""
        torch = _torch()
        if self.d_payload is None:
            raise RuntimeError("matrix is not on the device; call to_device() first")
        dev = self.d_payload.device
        squeeze = vec.dim() == 1
        v = _check.vectors("vec", vec, cols=self.cols, device=dev, dtype=torch.int8)
        n = v.shape[0]
        if out is None:
            out = torch.empty((n, self.rows), dtype=torch.int32, device=dev)
        _check.output("out", out, shape=(n, self.rows), dtype=torch.int32, device=dev)
        call("matvans_decompress_and_multiply_nvec", _ptr(out), _ptr(v), n, v.stride(0),
             self.rows, self.cols, _ptr(self.d_cursors), self.min_value, self.G,
             _ptr(self.d_cdf_data), _ptr(self.d_ppf_data), _ptr(self.d_payload),
             self.payload_size, _stream(stream))
        return out[0] if squeeze else out

    def backward(self, g, s_w, alpha=None, out=None, workspace=None, stream=None):
        """dx[v] = alpha[v] * W^T (s_w * g[v]) re-decoded from the compressed streams,
        g [rows] or [n_vec, rows] fp32 (n_vec <= 8); bitwise reproducible."""
        torch = _torch()
        if self.d_payload is None:
            raise RuntimeError("matrix is not on the device; call to_device() first")
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
        call("matvans_decompress_backward", _ptr(out), _ptr(gg), n, gg.stride(0), _ptr(s_w),
             _ptr(alpha), self.rows, self.cols, _ptr(self.d_cursors), self.min_value, self.G,
             _ptr(self.d_cdf_data), _ptr(self.d_ppf_data), _ptr(self.d_payload),
             self.payload_size, _ptr(workspace), workspace.numel() * workspace.element_size(),
             _stream(stream))
        return out[0] if squeeze else out

    def decompress_and_mult(self, d_result8, d_result32, vector, d_delta, stream=None):
        """Stream-ordered composite: decode-matvec -> absmax -> requant; delta stays on
        the device (matvansWarpPerRow.cu:150-176)."""
        torch = _torch()
        s = self.dev_struct()
        dev = self.d_payload.device
        _check.tensor("d_result8", d_result8, dtype=torch.int8, device=dev, min_numel=self.rows)
        _check.tensor("d_result32", d_result32, dtype=torch.int32, device=dev,
                      min_numel=self.rows)
        _check.tensor("vector", vector, dtype=torch.int8, device=dev, min_numel=self.cols)
        _check.tensor("d_delta", d_delta, dtype=torch.float32, device=dev, min_numel=1)
        call("matvans_decompress_and_mult", ctypes.byref(s), _ptr(d_result8), _ptr(d_result32),
             _ptr(vector), _ptr(d_delta), _stream(stream))

    def decompressAndMult(self, d_result8, d_result32, vector, v_delta=1.0):
        """Reference-shaped call (matvansWarpPerRow.cu:147): returns the new v_delta as a
        host float (synchronises, as the reference's thrust reduction does).  The input
        v_delta is ignored, as in the reference."""
        torch = _torch()
        d = torch.empty(1, dtype=torch.float32, device=d_result32.device)
        self.decompress_and_mult(d_result8, d_result32, vector, d)
        return float(d.item())


def _ppf_from_cdf(cdf16):
    c8 = (np.asarray(cdf16).astype(np.int64) & 0xFF)[:-1]
    ppf = np.zeros(256, np.uint8)
    for q in range(256):
        ok = np.nonzero(c8 <= q)[0]
        ppf[q] = ok.max() if len(ok) else 0
    return ppf


def read_bin(path):
    """Parse a .bin v0 file with the library reader.
    Returns (header dict, vector int8, [CompressedMatrix])."""
    h = ctypes.c_void_p()
    call("matvans_bin_read", str(path).encode(), ctypes.byref(h))
    try:
        num, rh, mwc, lv = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        vp = ctypes.c_void_p()
        call("matvans_bin_header", h, ctypes.byref(num), ctypes.byref(rh), ctypes.byref(mwc),
             ctypes.byref(vp), ctypes.byref(lv))
        vec = np.ctypeslib.as_array(ctypes.cast(vp, ctypes.POINTER(ctypes.c_int8)),
                                    (lv.value,)).copy() if lv.value else np.zeros(0, np.int8)
        mats = []
        for k in range(num.value):
            m = HostMatrix()
            call("matvans_bin_matrix", h, k, ctypes.byref(m))

            def arr(p, ct, n):
                return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), (n,)).copy()
            mats.append(CompressedMatrix(
                m.rows, m.cols, m.grid_spacing, arr(m.cursors, ctypes.c_uint32, m.rows),
                m.min_value, m.G, arr(m.cdf, ctypes.c_uint8, m.G + 1),
                arr(m.ppf, ctypes.c_uint8, 256),
                arr(m.payload, ctypes.c_uint16, m.payload_size) if m.payload_size
                else np.zeros(0, np.uint16)))
        hdr = {"num_matrices": num.value, "result_hash": rh.value, "max_word_count": mwc.value}
        return hdr, vec, mats
    finally:
        lib().matvans_bin_free(h)


def read_raw(path):
    """Parse an uncompressed baseline file (raw-matrices_*.bin) with the library reader
    (baseline_matVec_custom.cu:371-393,480-508).  Returns (header dict, vector int8,
    [int8 matrices [rows, cols]])."""
    h = ctypes.c_void_p()
    call("matvans_raw_read", str(path).encode(), ctypes.byref(h))
    try:
        num, rh, lv = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        vp = ctypes.c_void_p()
        call("matvans_raw_header", h, ctypes.byref(num), ctypes.byref(rh), ctypes.byref(vp),
             ctypes.byref(lv))
        vec = (np.ctypeslib.as_array(ctypes.cast(vp, ctypes.POINTER(ctypes.c_int8)),
                                     (lv.value,)).copy() if lv.value else np.zeros(0, np.int8))
        mats = []
        for k in range(num.value):
            r, c, gs, dp = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_float(), ctypes.c_void_p()
            call("matvans_raw_matrix", h, k, ctypes.byref(r), ctypes.byref(c), ctypes.byref(gs),
                 ctypes.byref(dp))
            n = r.value * c.value
            a = (np.ctypeslib.as_array(ctypes.cast(dp, ctypes.POINTER(ctypes.c_int8)), (n,)).copy()
                 if n else np.zeros(0, np.int8))
            mats.append(a.reshape(r.value, c.value))
        return {"num_matrices": num.value, "result_hash": rh.value}, vec, mats
    finally:
        lib().matvans_raw_free(h)


def write_bin(path, result_hash, vector, matrices):
    """Write a .bin v0 file with the library writer."""
    vec = np.ascontiguousarray(vector, np.int8)
    keep = []
    arr = (HostMatrix * max(1, len(matrices)))()
    for i, m in enumerate(matrices):
        cdf = np.ascontiguousarray(m.cdf_data, np.uint8)
        keep += [m.cursors, cdf, m.ppf_data, m.payload]
        arr[i] = HostMatrix(m.rows, m.cols, m.grid_spacing, m.min_value, len(cdf) - 1, 0,
                            m.payload_size, m.cursors.ctypes.data, cdf.ctypes.data,
                            m.ppf_data.ctypes.data, m.payload.ctypes.data)
    call("matvans_bin_write", str(path).encode(), int(result_hash), _np_ptr(vec), len(vec),
         len(matrices), arr)


# ------------------------------------------------------------------- device ops
def decmpressAndMultiply(dst, vec, rows, cols, grid_spacing, cursors, min_value, G, cdf_data,
                         ppf_data, payload, payload_size, stream=None):
    """The reference kernel's argument list (matvansWarpPerRow-kernel.cu:46-50) on torch
    device tensors; launches the fused HIP kernel on ``stream``."""
    torch = _torch()
    dev = payload.device
    _check.tensor("dst", dst, dtype=torch.int32, device=dev, min_numel=int(rows))
    _check.tensor("vec", vec, dtype=torch.int8, device=dev, min_numel=int(cols))
    _check.tensor("cursors", cursors, dtype=None, dtypes=(torch.int32, torch.uint32),
                  device=dev, min_numel=int(rows))
    _check.tensor("cdf_data", cdf_data, dtype=torch.uint8, device=dev, min_numel=int(G) + 1)
    _check.tensor("ppf_data", ppf_data, dtype=torch.uint8, device=dev, min_numel=256)
    _check.tensor("payload", payload, dtype=None, dtypes=(torch.int16, torch.uint16),
                  device=dev, min_numel=int(payload_size))
    call("matvans_decompress_and_multiply", _ptr(dst), _ptr(vec), int(rows), int(cols),
         float(grid_spacing), _ptr(cursors), int(min_value), int(G), _ptr(cdf_data),
         _ptr(ppf_data), _ptr(payload), int(payload_size), _stream(stream))


def decompress_and_multiply(matrix: CompressedMatrix, vec, out=None, stream=None):
    return matrix.multiply(vec, out, stream)


def int8_matvec(mat, vecs, out=None, stream=None):
    """Uncompressed int8 comparator: int32 mat @ vecs^T, vecs [n_vec, cols] or [cols]."""
    torch = _torch()
    _check.tensor("mat", mat, dtype=torch.int8)
    if mat.dim() != 2 or not mat.is_contiguous():
        raise ValueError("mat: expected a contiguous [rows, cols] int8 matrix")
    rows, cols = mat.shape
    squeeze = vecs.dim() == 1
    v = _check.vectors("vecs", vecs, cols=cols, device=mat.device, dtype=torch.int8)
    if out is None:
        out = torch.empty((v.shape[0], rows), dtype=torch.int32, device=mat.device)
    _check.output("out", out, shape=(v.shape[0], rows), dtype=torch.int32, device=mat.device)
    call("matvans_int8_matvec", _ptr(out), _ptr(mat), _ptr(v), v.shape[0], v.stride(0), rows,
         cols, _stream(stream))
    return out[0] if squeeze else out


def f32_gemv(mat, x, alpha=1.0, out=None, stream=None):
    """fp32 comparator: alpha * mat @ x."""
    torch = _torch()
    _check.tensor("mat", mat, dtype=torch.float32)
    if mat.dim() != 2 or not mat.is_contiguous():
        raise ValueError("mat: expected a contiguous [rows, cols] fp32 matrix")
    rows, cols = mat.shape
    _check.tensor("x", x, dtype=torch.float32, device=mat.device, shape=(cols,))
    if out is None:
        out = torch.empty(rows, dtype=torch.float32, device=mat.device)
    _check.tensor("out", out, dtype=torch.float32, device=mat.device, shape=(rows,))
    call("matvans_f32_gemv", _ptr(out), _ptr(mat), _ptr(x), rows, cols, float(alpha),
         _stream(stream))
    return out


def absmax_requant(acc, clamp=False, out8=None, delta=None, stream=None):
    """(int8 requantised, fp32 delta tensor[1]) of an int32 accumulator vector."""
    torch = _torch()
    _check.tensor("acc", acc, dtype=torch.int32)
    if not acc.is_contiguous():
        raise ValueError("acc: must be contiguous")
    n = acc.numel()
    if out8 is None:
        out8 = torch.empty(n, dtype=torch.int8, device=acc.device)
    if delta is None:
        delta = torch.empty(1, dtype=torch.float32, device=acc.device)
    _check.tensor("out8", out8, dtype=torch.int8, device=acc.device, min_numel=n)
    _check.tensor("delta", delta, dtype=torch.float32, device=acc.device, min_numel=1)
    call("matvans_absmax_requant", _ptr(acc), n, _ptr(out8), _ptr(delta), int(bool(clamp)),
         _stream(stream))
    return out8, delta


def f32_normalize_round(y, delta=None, stream=None):
    """The fp32 chain's epilogue, in place: y = roundf(y / delta), delta = max|y| / 127
    (round half away from zero).  Returns (y, fp32 delta tensor[1])."""
    torch = _torch()
    _check.tensor("y", y, dtype=torch.float32)
    if not y.is_contiguous():
        raise ValueError("y: must be contiguous")
    if y.numel() == 0:
        raise ValueError("y: empty")
    if delta is None:
        delta = torch.empty(1, dtype=torch.float32, device=y.device)
    _check.tensor("delta", delta, dtype=torch.float32, device=y.device, min_numel=1)
    call("matvans_f32_normalize_round", _ptr(y), y.numel(), _ptr(delta), _stream(stream))
    return y, delta


def scale_output(acc, s_w, s_x, out=None, stream=None):
    """y[v, r] = s_x[v] * s_w[r] * float(acc[v, r])."""
    torch = _torch()
    _check.tensor("acc", acc, dtype=torch.int32)
    a = acc if acc.dim() == 2 else acc.unsqueeze(0)
    if not a.is_contiguous():
        raise ValueError("acc: must be contiguous")
    n_vec, rows = a.shape
    _check.tensor("s_w", s_w, dtype=torch.float32, device=acc.device, shape=(rows,))
    _check.tensor("s_x", s_x, dtype=torch.float32, device=acc.device, shape=(n_vec,))
    if out is None:
        out = torch.empty((n_vec, rows), dtype=torch.float32, device=acc.device)
    _check.output("out", out, shape=(n_vec, rows), dtype=torch.float32, device=acc.device)
    call("matvans_scale_output", _ptr(a), rows, n_vec, _ptr(s_w), _ptr(s_x), _ptr(out),
         _stream(stream))
    return out if acc.dim() == 2 else out[0]


def quantize_vectors(x, stream=None):
    """Per-vector activation quantiser: (int8 xq [n, cols], fp32 s_x [n]) with
    s_x = max|x| / 127 and xq = clamp(rint(x / s_x), -127, 127)."""
    torch = _torch()
    _check.tensor("x", x, dtype=torch.float32)
    xx = (x.unsqueeze(0) if x.dim() == 1 else x).contiguous()
    n, cols = xx.shape
    xq = torch.empty((n, cols), dtype=torch.int8, device=xx.device)
    s_x = torch.empty(n, dtype=torch.float32, device=xx.device)
    call("matvans_quantize_vectors", _ptr(xx), n, xx.stride(0), cols, _ptr(xq), xq.stride(0),
         _ptr(s_x), _stream(stream))
    return xq, s_x


def _compressed_linear_function():
    torch = _torch()

    class CompressedLinearFunction(torch.autograd.Function):
        """y = s_x * s_w * (Wq @ quantize(x)) with W frozen in its compressed form.

        Backward uses the straight-through estimator for the activation quantiser
        (d quantize(x) / dx = 1 / s_x, s_x treated as a constant), so
        dL/dx = Wq^T (s_w * dL/dy), computed from the compressed streams."""

        @staticmethod
        def forward(ctx, x, matrix, s_w):
            shape = x.shape
            xx = x.reshape(-1, matrix.cols).float().contiguous()
            if xx.shape[0] > 8:
                raise ValueError("at most 8 vectors per call")
            xq, s_x = quantize_vectors(xx)
            acc = matrix.multiply(xq)
            y = scale_output(acc, s_w, s_x)
            ctx.matrix = matrix
            ctx.in_shape = shape
            ctx.save_for_backward(s_w)
            return y.reshape(*shape[:-1], matrix.rows)

        @staticmethod
        def backward(ctx, gy):
            (s_w,) = ctx.saved_tensors
            g = gy.reshape(-1, ctx.matrix.rows).float().contiguous()
            dx = ctx.matrix.backward(g, s_w)
            return dx.reshape(ctx.in_shape), None, None

    return CompressedLinearFunction


_CLF = None


def compressed_linear(x, matrix, s_w):
    """Autograd-aware compressed linear layer (see CompressedLinearFunction)."""
    global _CLF
    if _CLF is None:
        _CLF = _compressed_linear_function()
    return _CLF.apply(x, matrix, s_w)


def run_chain(matrices, vector, stream=None):
    """The reference benchmark chain v_i = requant(W_i v_{i-1}) on the device
    (matvansWarpPerRow.cu:320-350).  ``matrices`` are on the device; ``vector`` is an
    int8 device tensor.  Returns the final int8 device vector (one host sync at the end
    is left to the caller)."""
    torch = _torch()
    dev = vector.device
    max_rows = max(m.rows for m in matrices)
    r32 = torch.empty(max_rows, dtype=torch.int32, device=dev)
    bufs = [torch.empty(max(max_rows, vector.numel()), dtype=torch.int8, device=dev)
            for _ in range(2)]
    delta = torch.empty(1, dtype=torch.float32, device=dev)
    cur = vector
    for k, m in enumerate(matrices):
        out = bufs[k % 2]
        m.decompress_and_mult(out, r32, cur, delta, stream)
        cur = out[:m.rows]
    return cur
