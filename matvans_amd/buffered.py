"""The reference's family-B path ("buffered": quad rANS, layouts ans-warp-rows and
ans-warp-stripes) over the C ABI of ``lib/libmatvans.so``.

  BufferedMatrices.read(path)          safetensors file of buffered/encoder.py
  BufferedMatrices.encode(mats, ...)   compress_all_matrices (encoder.py:474-510)
  .to_device() / .mat_vec(k, ...)      mat_vec_mul_coalesced (ans-warp-rows.cu:35-56)
  .run_chain()                         the matrix chain of main (ans-warp-rows.cu:120-186)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check
from .capi import call, lib

__all__ = ["BufferedMatrices", "read_safetensors"]

_DT = {"U32": np.uint32, "I8": np.int8, "F32": np.float32, "U8": np.uint8, "I32": np.int32,
       "F64": np.float64, "I64": np.int64, "U64": np.uint64, "I16": np.int16, "U16": np.uint16}


def _np_ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def read_safetensors(path):
    """All tensors of a safetensors file as numpy arrays (library reader)."""
    h = ctypes.c_void_p()
    call("matvans_st_read", str(path).encode(), ctypes.byref(h))
    try:
        import json
        import struct
        with open(path, "rb") as f:
            (n,) = struct.unpack("<Q", f.read(8))
            names = [k for k in json.loads(f.read(n)) if k != "__metadata__"]
        out = {}
        for name in names:
            dt = ctypes.c_char_p()
            nd = ctypes.c_uint32()
            shape = (ctypes.c_uint64 * 4)()
            data = ctypes.c_void_p()
            nb = ctypes.c_uint64()
            call("matvans_st_tensor", h, name.encode(), ctypes.byref(dt), ctypes.byref(nd),
                 shape, ctypes.byref(data), ctypes.byref(nb))
            dtype = np.dtype(_DT[dt.value.decode()])
            shp = tuple(int(shape[i]) for i in range(nd.value))
            buf = (ctypes.c_uint8 * nb.value).from_address(data.value) if nb.value else b""
            out[name] = np.frombuffer(bytes(buf), dtype=dtype).reshape(shp).copy()
        return out
    finally:
        lib().matvans_st_free(h)


class BufferedMatrices:
    """num matrices of dim x dim (or rows x cols) in one family-B layout."""

    def __init__(self, layout, compressed, all_warp_offsets, all_ppfs, all_ppfs_cum,
                 all_ppfs_prob, scales, initial_vector=None, hashes=None, shapes=None):
        assert layout in ("rows", "stripes")
        self.layout = layout
        self.compressed = np.ascontiguousarray(compressed, np.uint32)
        self.offsets = [np.ascontiguousarray(o, np.uint32) for o in all_warp_offsets]
        self.ppfs = np.ascontiguousarray(all_ppfs, np.uint32)
        self.cums = np.ascontiguousarray(all_ppfs_cum, np.uint32)
        self.probs = np.ascontiguousarray(all_ppfs_prob, np.uint32)
        self.scales = np.ascontiguousarray(scales, np.float32)
        self.initial_vector = initial_vector
        self.hashes = hashes
        n = len(self.scales)
        if shapes is None:
            dim = len(initial_vector)
            shapes = [(dim, dim)] * n
        self.shapes = [tuple(int(x) for x in s) for s in shapes]
        self.d = None

    @property
    def num(self):
        return len(self.scales)

    def end_offset(self, k):
        """End of matrix k's data in 16-byte units."""
        return int(self.offsets[k + 1][0]) if k + 1 < self.num else len(self.compressed) // 4

    @classmethod
    def read(cls, path):
        t = read_safetensors(path)
        dim = len(t["initial_vector"])
        offs = t["all_warp_offsets"]
        layout = "rows" if offs.shape[1] == dim else "stripes"
        return cls(layout, t["compressed"], list(offs), t["all_ppfs"], t["all_ppfs_cum"],
                   t["all_ppfs_prob"], t["scales"], t["initial_vector"], t["hashes"])

    @classmethod
    def encode(cls, matrices, layout="rows", scales=None, initial_vector=None, hashes=None):
        """compress_all_matrices (encoder.py:474-495) with the library model and encoder."""
        li = 0 if layout == "rows" else 1
        parts, offs, ppfs, cums, probs, size = [], [], [], [], [], 0
        for m in matrices:
            m = np.ascontiguousarray(m, np.int8)
            rows, cols = m.shape
            mn, ns = ctypes.c_int32(), ctypes.c_uint32()
            cdf = np.zeros(257, np.uint8)
            p, c, f = (np.zeros(32, np.uint32) for _ in range(3))
            call("matvans_b_model", _np_ptr(m), m.size, ctypes.byref(mn), ctypes.byref(ns),
                 _np_ptr(cdf), _np_ptr(p), _np_ptr(c), _np_ptr(f))
            cap = int(lib().matvans_b_encode_bound(rows, cols, li))
            out = np.empty(cap, np.uint32)
            o = np.empty(rows if li == 0 else rows // 32, np.uint32)
            nw = ctypes.c_uint64()
            call("matvans_b_encode", _np_ptr(m), rows, cols, li, mn.value, _np_ptr(cdf), ns.value,
                 size, _np_ptr(out), cap, ctypes.byref(nw), _np_ptr(o))
            parts.append(out[:nw.value].copy())
            offs.append(o)
            ppfs.append(p)
            cums.append(c)
            probs.append(f)
            size += nw.value // 4
        n = len(parts)
        return cls(layout, np.concatenate(parts), offs, np.stack(ppfs), np.stack(cums),
                   np.stack(probs), np.ones(n, np.float32) if scales is None else scales,
                   initial_vector, hashes, shapes=[np.shape(m) for m in matrices])

    def to_device(self, device="cuda"):
        import torch
        dev = torch.device(device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to(dev)
        self.d = {"compressed": t(self.compressed), "offsets": [t(o) for o in self.offsets],
                  "ppf": t(self.ppfs), "cum": t(self.cums), "prob": t(self.probs)}
        return self

    def mat_vec(self, k, vecs, out8=None, acc32=None, scale=None, stream=None):
        """Matrix k times 1..8 int8 vectors [n_vec, cols] (or [cols]) on the device.
        Writes out8 = clamp(rint(float(acc) * scale)) and/or acc32."""
        import torch
        if self.d is None:
            raise RuntimeError("call to_device() first")
        if not 0 <= k < self.num:
            raise ValueError(f"k: matrix index {k} outside 0..{self.num - 1}")
        rows, cols = self.shapes[k]
        dev = self.d["compressed"].device
        v = _check.vectors("vecs", vecs, cols=cols, device=dev, dtype=torch.int8)
        n = v.shape[0]
        if out8 is None and acc32 is None:
            acc32 = torch.empty((n, rows), dtype=torch.int32, device=dev)
        if acc32 is not None:
            _check.output("acc32", acc32, shape=(n, rows), dtype=torch.int32, device=dev)
        if out8 is not None:
            _check.output("out8", out8, shape=(n, rows), dtype=torch.int8, device=dev)
        fn = "matvans_b_rows_matvec" if self.layout == "rows" else "matvans_b_stripes_matvec"
        rowp = lambda t: ctypes.c_void_p(t.data_ptr() + 128 * k)
        call(fn, _ptr(self.d["compressed"]), self.compressed.size, _ptr(self.d["offsets"][k]),
             self.end_offset(k), rowp(self.d["ppf"]), rowp(self.d["cum"]), rowp(self.d["prob"]),
             _ptr(v), n, v.stride(0), cols, rows,
             float(self.scales[k] if scale is None else scale),
             _ptr(out8) if out8 is not None else None, _ptr(acc32) if acc32 is not None else None,
             _stream(stream))
        return acc32 if acc32 is not None else out8

    def run_chain(self, vector=None, stream=None):
        """The reference's benchmark chain (ans-warp-rows.cu:143-176): each matrix maps the
        int8 vector to clamp(rint(W v * scale_k)).  Returns the final int8 vector (device)."""
        import torch
        v = vector if vector is not None else torch.from_numpy(self.initial_vector).cuda()
        for k in range(self.num):
            out = torch.empty(self.shapes[k][0], dtype=torch.int8, device=v.device)
            self.mat_vec(k, v, out8=out, stream=stream)
            v = out
        return v
