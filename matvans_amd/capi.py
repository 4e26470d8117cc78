"""ctypes binding of ``libmatvans.so`` (declared in ``include/matvans.h``).

The library is loaded from ``matvans_amd/lib/libmatvans.so``.  There is no fallback:
if the library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmatvans.so")

P = ctypes.c_void_p
u8, i8, u16, u32, u64, i32, f32 = (ctypes.c_uint8, ctypes.c_int8, ctypes.c_uint16,
                                   ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float)


class MatvansError(RuntimeError):
    """A matvans_* call returned a nonzero status."""

    def __init__(self, fn, status, msg):
        super().__init__(f"{fn} -> status {status}: {msg}")
        self.status = status


class DevMatrix(ctypes.Structure):
    """matvans_dev_matrix."""
    _fields_ = [("rows", u32), ("cols", u32), ("grid_spacing", f32), ("min_value", i8),
                ("G", u8), ("reserved", u16), ("payload_size", u32), ("cursors", P),
                ("cdf", P), ("ppf", P), ("payload", P)]


class HostMatrix(ctypes.Structure):
    """matvans_host_matrix."""
    _fields_ = [("rows", u32), ("cols", u32), ("grid_spacing", f32), ("min_value", i8),
                ("G", u8), ("reserved", u16), ("payload_size", u32), ("cursors", P),
                ("cdf", P), ("ppf", P), ("payload", P)]


class V1Host(ctypes.Structure):
    """matvans_v1_host."""
    _fields_ = [("rows", u32), ("cols", u32), ("grid_spacing", f32), ("streams", u32),
                ("precision", u32), ("table_mode", u32), ("coder", u32), ("min_value", i32),
                ("G", u32), ("payload_size", u32), ("cursors", P), ("cdf", P), ("freq", P),
                ("payload", P), ("n_outliers", u32), ("out_start", P), ("out_rec", P)]


class V1DevMatrix(ctypes.Structure):
    """matvans_v1_dev_matrix."""
    _fields_ = [("rows", u32), ("cols", u32), ("streams", u32), ("precision", u32),
                ("table_mode", u32), ("coder", u32), ("min_value", i32), ("G", u32),
                ("payload_size", u32), ("cursors", P), ("lut", P), ("freq", P), ("payload", P),
                ("n_outliers", u32), ("out_start", P), ("out_rec", P), ("out_col_start", P),
                ("out_col_rec", P)]


class BatchItem(ctypes.Structure):
    """matvans_batch_item."""
    _fields_ = [("format", u32), ("rows", u32), ("cols", u32), ("n_vec", u32),
                ("vec_stride", u32), ("vecs", P), ("dst", P), ("cursors", P), ("payload", P),
                ("payload_size", u32), ("min_value", i32), ("G", u32), ("cdf", P), ("ppf", P),
                ("streams", u32), ("lut", P), ("compressed", P), ("compressed_words", u64),
                ("warp_offsets", P), ("end_offset", u32), ("b_ppf", P), ("b_cum", P),
                ("b_prob", P), ("scale", f32), ("acc32", P), ("out8", P), ("out_start", P),
                ("out_rec", P)]


class BatchBwdItem(ctypes.Structure):
    """matvans_batch_bwd_item."""
    _fields_ = [("format", u32), ("rows", u32), ("cols", u32), ("n_vec", u32),
                ("g_stride", u32), ("g", P), ("s_w", P), ("alpha", P), ("dx", P),
                ("cursors", P), ("payload", P), ("payload_size", u32), ("min_value", i32),
                ("G", u32), ("cdf", P), ("ppf", P), ("streams", u32), ("lut", P),
                ("out_col_start", P), ("out_col_rec", P)]


class BatchConfig(ctypes.Structure):
    """matvans_batch_config."""
    _fields_ = [("units_per_tile", u32), ("order", u32), ("blocks_per_cu", u32),
                ("schedule", u32)]


# name -> (restype, argtypes); status-returning functions use i32
SIGNATURES = {
    "matvans_last_error": (ctypes.c_char_p, []),
    "matvans_version": (i32, []),
    "matvans_set_decode_variant": (i32, [i32]),
    "matvans_get_decode_variant": (i32, []),
    "matvans_decompress_and_multiply": (i32, [P, P, u32, u32, f32, P, i8, u8, P, P, P, u32, P]),
    "matvans_decompress_and_multiply_nvec": (i32, [P, P, u32, u32, u32, u32, P, i8, u8, P, P,
                                                   P, u32, P]),
    "matvans_decompress_and_mult": (i32, [ctypes.POINTER(DevMatrix), P, P, P, P, P]),
    "matvans_absmax_requant": (i32, [P, u32, P, P, i32, P]),
    "matvans_f32_normalize_round": (i32, [P, u32, P, P]),
    "matvans_scale_output": (i32, [P, u32, u32, P, P, P, P]),
    "matvans_int8_matvec": (i32, [P, P, P, u32, u32, u32, u32, P]),
    "matvans_f32_gemv": (i32, [P, P, P, u32, u32, f32, P]),
    "matvans_backward_workspace_size": (ctypes.c_size_t, [u32, u32, u32]),
    "matvans_decompress_backward": (i32, [P, P, u32, u32, P, P, u32, u32, P, i8, u8, P, P, P,
                                          u32, P, ctypes.c_size_t, P]),
    "matvans_quantize_vectors": (i32, [P, u32, u32, u32, P, u32, P, P]),
    "matvans_fxhash": (u32, [P, u64]),
    "matvans_bin_read": (i32, [ctypes.c_char_p, ctypes.POINTER(P)]),
    "matvans_bin_free": (None, [P]),
    "matvans_raw_read": (i32, [ctypes.c_char_p, ctypes.POINTER(P)]),
    "matvans_raw_free": (None, [P]),
    "matvans_raw_header": (i32, [P, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(P),
                                 ctypes.POINTER(u32)]),
    "matvans_raw_matrix": (i32, [P, u32, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                 ctypes.POINTER(f32), ctypes.POINTER(P)]),
    "matvans_bin_header": (i32, [P, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                 ctypes.POINTER(u32), ctypes.POINTER(P), ctypes.POINTER(u32)]),
    "matvans_bin_matrix": (i32, [P, u32, ctypes.POINTER(HostMatrix)]),
    "matvans_bin_write": (i32, [ctypes.c_char_p, u32, P, u32, u32,
                                ctypes.POINTER(HostMatrix)]),
    "matvans_build_model": (i32, [P, u64, ctypes.POINTER(i8), ctypes.POINTER(u32), P, P]),
    "matvans_encode_bound": (u64, [u32, u32]),
    "matvans_encode_rows": (i32, [P, u32, u32, i8, u32, P, u32, P, P, u64,
                                  ctypes.POINTER(u64)]),
    "matvans_v1_model": (i32, [P, u32, u32, u32, u32, ctypes.POINTER(i32), ctypes.POINTER(u32),
                               P, P]),
    "matvans_v1_encode_bound": (u64, [u32, u32, u32, u32]),
    "matvans_v1_encode": (i32, [P, u32, u32, u32, u32, u32, u32, i32, u32, P, P, u32, P, P,
                                u64, ctypes.POINTER(u64)]),
    "matvans_v1_lut": (i32, [P, u32, i32, u32, P]),
    "matvans_v1_outlier_band": (i32, [P, u64, ctypes.POINTER(i32)]),
    "matvans_v1_split_outliers": (i32, [P, u32, u32, i32, P, P, P, u64, ctypes.POINTER(u64)]),
    "matvans_v1_read": (i32, [ctypes.c_char_p, ctypes.POINTER(P)]),
    "matvans_v1_free": (None, [P]),
    "matvans_v1_header": (i32, [P, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(P),
                                ctypes.POINTER(u32)]),
    "matvans_v1_matrix": (i32, [P, u32, ctypes.POINTER(V1Host)]),
    "matvans_v1_write": (i32, [ctypes.c_char_p, u32, P, u32, u32, ctypes.POINTER(V1Host)]),
    "matvans_v1_decompress_and_multiply": (i32, [ctypes.POINTER(V1DevMatrix), P, P, u32, u32, P]),
    "matvans_matrix_upload": (i32, [ctypes.POINTER(HostMatrix), ctypes.POINTER(P),
                                    ctypes.POINTER(DevMatrix)]),
    "matvans_v1_matrix_upload": (i32, [ctypes.POINTER(V1Host), ctypes.POINTER(P),
                                       ctypes.POINTER(V1DevMatrix)]),
    "matvans_upload_free": (None, [P]),
    "matvans_v1_decompress_backward": (i32, [ctypes.POINTER(V1DevMatrix), P, P, u32, u32, P, P,
                                             P, ctypes.c_size_t, P]),
    "matvans_v1_decompress_backward_noalloc": (i32, [ctypes.POINTER(V1DevMatrix), P, P, u32, u32,
                                                     P, P, P]),
    "matvans_b_rows_matvec": (i32, [P, u64, P, u32, P, P, P, P, u32, u32, u32, u32, f32, P, P, P]),
    "matvans_b_stripes_matvec": (i32, [P, u64, P, u32, P, P, P, P, u32, u32, u32, u32, f32, P, P,
                                       P]),
    "matvans_b_model": (i32, [P, u64, ctypes.POINTER(i32), ctypes.POINTER(u32), P, P, P, P]),
    "matvans_b_encode_bound": (u64, [u32, u32, i32]),
    "matvans_b_encode": (i32, [P, u32, u32, i32, i32, P, u32, u32, P, u64, ctypes.POINTER(u64),
                               P]),
    "matvans_batch_create": (i32, [ctypes.POINTER(BatchItem), u32, ctypes.POINTER(P)]),
    "matvans_batch_destroy": (None, [P]),
    "matvans_batch_create_backward": (i32, [ctypes.POINTER(BatchBwdItem), u32,
                                            ctypes.POINTER(P)]),
    "matvans_batch_configure": (i32, [P, ctypes.POINTER(BatchConfig)]),
    "matvans_batch_get_config": (i32, [P, ctypes.POINTER(BatchConfig)]),
    "matvans_batch_launch": (i32, [P, P]),
    "matvans_batch_autotune": (i32, [P, P, u32, ctypes.POINTER(BatchConfig),
                                     ctypes.POINTER(f32)]),
    "matvans_st_read": (i32, [ctypes.c_char_p, ctypes.POINTER(P)]),
    "matvans_st_free": (None, [P]),
    "matvans_st_tensor": (i32, [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
                                ctypes.POINTER(u32), P, ctypes.POINTER(P), ctypes.POINTER(u64)]),
}

_LIB = None


_DEFAULT_LIB_PATH = LIB_PATH


def _check_stamp():
    """The in-tree library must have been built from the sources beside it: build() writes
    a digest of the sources next to it, and a library whose digest differs (sources edited
    or reverted without a rebuild) is refused rather than run."""
    from . import build as B
    try:
        with open(B.STAMP) as fh:
            built = fh.read().strip()
    except FileNotFoundError:
        raise RuntimeError(f"{LIB_PATH} has no source stamp ({B.STAMP}); rebuild with "
                           "python -m matvans_amd.build") from None
    if built != B.source_digest():
        raise RuntimeError(f"{LIB_PATH} was built from other sources than matvans_amd/csrc "
                           "and include/matvans.h hold now; rebuild with "
                           "python -m matvans_amd.build")


def _bind_hip_runtime():
    """Map the HIP runtime that the library's calls must share with torch before the
    library itself.  libmatvans.so needs libamdhip64.so.7; torch ships its own copy under
    torch/lib.  If the library is loaded first, the dynamic loader maps ROCm's copy (and
    its libhsa-runtime64), and a later ``import torch`` maps torch's copy beside it: two
    HIP and two HSA runtimes in one process, each with its own device state, which has
    crashed GPU calls.  Importing torch first maps torch's runtime, and the library's
    dependency then resolves to it (same soname), so one runtime serves both.  Without
    torch the library uses ROCm's runtime alone, which is also a single runtime."""
    try:
        import torch  # noqa: F401  (imported for its side effect on the loader)
    except ImportError:
        pass


def hip_runtimes_mapped() -> list[str]:
    """Paths of every libamdhip64 mapped into this process (Linux /proc/self/maps)."""
    seen = []
    try:
        with open("/proc/self/maps") as fh:
            for line in fh:
                parts = line.split()
                if len(parts) >= 6 and "libamdhip64" in parts[-1]:
                    path = os.path.realpath(parts[-1])
                    if path not in seen:
                        seen.append(path)
    except OSError:
        pass
    return seen


def lib():
    """The loaded library; raises if it is not built or was built from other sources."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run python -m matvans_amd.build")
        if LIB_PATH == _DEFAULT_LIB_PATH:
            _check_stamp()
        _bind_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        hips = hip_runtimes_mapped()
        if len(hips) > 1:
            raise RuntimeError("two HIP runtimes are mapped in this process (" +
                               ", ".join(hips) + "); load torch before any other HIP "
                               "library so that libmatvans.so shares torch's runtime")
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def last_error() -> str:
    return lib().matvans_last_error().decode()


def call(name, *args):
    """Call a status-returning entry point; raise MatvansError on failure."""
    st = getattr(lib(), name)(*args)
    if st != 0:
        raise MatvansError(name, st, last_error())
    return st
