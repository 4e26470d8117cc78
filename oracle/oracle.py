"""CPU oracle for the WarpPerRow rANS codec and its fused decode-matvec.

TEST INFRASTRUCTURE ONLY.  Nothing under ``oracle/`` is part of the product: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only to check results.  The product path (``matvans_amd``) never
calls into it and fails loudly when its HIP library is missing.

This module is a numpy restatement of the reference algorithm.  Each function cites
the reference file:line it follows.  ``oracle/oracle.c`` restates the same decoder
and encoder in plain C for sizes where numpy is too slow; ``tests/test_oracle.py``
pins both against the golden fixtures written by the reference generator
(``tests/golden/``), against each other, and against closed forms (round trips,
brute-force integer matvec, the FxHash definition).

Pinned by:
  fxhash ................ golden ``result_hash`` (header) via ``reference_chain``
  gen_reference_matrices  golden raw-matrices file (byte-exact)
  reference_chain ....... golden ``result_hash`` (both fixtures)
  create_entropy_model .. golden cdf (byte-exact, both fixtures)
  create_ppf ............ golden ppf (byte-exact)
  encode_matrix ......... golden payload outside padding, golden cursors
  decode_matrix ......... decode(golden) == golden raw matrix; round trip
  read_bin / write_bin .. byte round trip of the golden files
  requant_chain_f32 ..... closed form (IEEE fp32, round-half-even), see tests
"""
from __future__ import annotations

import gzip
import struct

import numpy as np

PRECISION = 8          # create_matrices_WarpPerRow.py:33
WORD_BITS = 16         # AnsCoder(precision, 16, ...) :372
LANES = 32             # streams per row, one per warp lane (:363)
GRANULE_WORDS = 8      # u16 words per lane per 16-byte granule (:405)
GROUP_COLS = 512       # 32 lanes x 16 columns (:365-366)
DEFAULT_SEED = 20250310  # :10


# --------------------------------------------------------------------------- hash
def fxhash(arr) -> int:
    """FxHash of an int8 array (create_matrices_WarpPerRow.py:58-70,
    matvansWarpPerRow.cu:31-42): h = rotl(h, 5) ^ byte; h *= 0x27220A95 (mod 2^32)."""
    a = np.asarray(arr)
    if a.dtype != np.int8:
        a = a.astype(np.int8)
    h = 0
    for b in a.view(np.uint8).ravel().tolist():
        h = ((h >> 27) | (h << 5)) & 0xFFFFFFFF
        h = ((h ^ b) * 0x27220A95) & 0xFFFFFFFF
    return h


# ---------------------------------------------------------------- data generation
def gen_reference_matrices(w: int, n: int, entropy: int, seed: int = DEFAULT_SEED):
    """Seeded square matrices on an ``entropy``-bit grid re-spread to +-127, plus the
    start vector (create_matrices_WarpPerRow.py:10, :35-50).
    Returns (quantized int8 [n,w,w], w_deltas f64 [n], vector int8 [w])."""
    np.random.seed(seed)
    sigma = 1 / np.sqrt(w)
    q = np.empty((n, w, w), dtype=np.int8)
    w_deltas = np.zeros(n)
    max_value = (2 ** (entropy - 1)) - 1
    for i in range(n):
        m = np.random.randn(w, w) * sigma
        w_delta = np.abs(m).max() / max_value
        w_deltas[i] = w_delta
        q[i, ...] = np.round(np.round(m / w_delta) * 127 / max_value).astype(np.int8)
    vector = np.round(np.random.randn(w) * 4).astype(np.int8)
    return q, w_deltas, vector


def reference_chain(mats, w_deltas, vector):
    """Float64 chain v_i = requant(W_i v_{i-1}) whose final hash is the ``.bin`` header's
    ``result_hash`` (create_matrices_WarpPerRow.py:72-92).  Returns (int8 vector, hash)."""
    v_delta = 1
    v = vector.copy()
    for i, mat in enumerate(mats):
        v32 = mat.astype(float) @ np.asarray(v).astype(float)
        vf = v32 * w_deltas[i] * v_delta
        v_delta = np.abs(vf).max() / 127
        v = np.round(vf / v_delta)
    out = np.array(v, dtype=np.int8)
    return out, fxhash(out)


def requant_f32(acc, clamp=False):
    """The reference device epilogue (matvansWarpPerRow.cu:51-72): a = float(max|acc|)
    (thrust transform_reduce; the max is taken on integer-valued floats), delta = a/127
    in fp32, y = (int8) rint(float(acc) / delta) in fp32 with round-half-even, no clamp.
    clamp=True is the int8 baseline's variant (baselines/baseline_int8/
    baseline_matVec_custom.cu:64-86: the max over fabs as float, __float2int_rn, then
    max(-128, min(127, .))).  Returns (int8 y, float32 delta)."""
    acc = np.asarray(acc, dtype=np.int32)
    a = np.float32(np.abs(acc.astype(np.int64)).max()) if acc.size else np.float32(0)
    delta = np.float32(a / np.float32(127))
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.rint(acc.astype(np.float32) / delta)
    if not np.all(np.isfinite(q)):
        return np.zeros_like(acc, dtype=np.int8), delta
    q = q.astype(np.int64)
    if clamp:
        q = np.clip(q, -128, 127)
    return q.astype(np.int8), delta


def normalize_round_f32(y):
    """The fp32 comparator chain's epilogue (baselines/cuBLAS_baseline_float/
    baseline_matVec_cuBLAS.cu:48-64, 129-135): a = max|y| (thrust transform_reduce of fabs
    with maximum), delta = a / 127 in fp32, y = roundf(y / delta) in fp32 (round half away
    from zero).  Returns (float32 y, float32 delta)."""
    y = np.asarray(y, dtype=np.float32)
    a = np.abs(y).max() if y.size else np.float32(0)
    delta = np.float32(np.float32(a) / np.float32(127))
    with np.errstate(divide="ignore", invalid="ignore"):
        q = (y / delta).astype(np.float32)
        t = np.trunc(q)
        f = q - t                       # exact: q and trunc(q) share the exponent range
        r = t + np.where(np.abs(f) >= np.float32(0.5), np.sign(q), np.float32(0))
    return r.astype(np.float32), delta


# ------------------------------------------------------------------ entropy model
def create_entropy_model(data, precision: int = PRECISION):
    """8-bit quantised model (create_matrices_WarpPerRow.py:131-187).

    Follows the reference step for step, including ``np.argsort(counts)[::-1]`` (whose
    tie order is numpy's) and ``np.round`` (half-even).  One deliberate difference: the
    pmf slot is ``int(value) - min_value + 1``.  The reference computes it with an int8
    ``value`` (:180-181), which wraps under numpy 2 promotion; the two agree whenever
    ``max - min <= 127`` or ``G + 1 == 256`` (every file the reference can write).
    Returns (min_value, grid_size G, cdf uint16[G+1], entropy, cross_entropy)."""
    data = np.asarray(data)
    min_value = int(data.min())
    max_value = int(data.max())
    if min_value == max_value:
        cdf = np.array([0, (1 << precision) - 1, 1 << precision], dtype=np.uint16)
        return min_value, 2, cdf, 0.0, float(-np.log1p(-1 / (1 << precision)) / np.log(2))
    grid_size = max_value - min_value + 1
    values, counts = np.unique(data, return_counts=True)
    order = np.argsort(counts)[::-1]
    values = values[order]
    counts = counts[order]
    scale = (1 << precision) / data.size
    qc = np.maximum(1, np.round(counts * scale).astype(np.uint16))
    excess = int(qc.sum()) - (1 << precision)
    if excess > 0:
        assert excess <= len(counts)
        while excess > 0:
            nred = int((qc > 1).sum())
            nr = min(excess, nred)
            assert nr > 0
            qc[:nr] -= 1
            excess -= nr
        qc[:excess] -= 1          # no-op (excess == 0), kept as in the reference (:166)
    elif excess < 0:
        assert -excess <= len(counts)
        qc[excess:] += 1
    assert int(qc.sum()) == 1 << precision
    entropy = float(np.log2(data.size) - (counts @ np.log2(counts)) / data.size)
    cross = float(precision - (counts @ np.log2(qc)) / data.size)
    pmf = np.zeros(grid_size + 1, dtype=np.uint16)
    for value, prob in zip(values.tolist(), qc.tolist()):
        pmf[value - min_value + 1] = prob
    cdf = np.cumsum(pmf, dtype=np.uint16)
    assert cdf[0] == 0 and cdf[-1] == 1 << precision
    return min_value, grid_size, cdf, entropy, cross


def create_ppf(cdf):
    """ppf[q] = largest s in [0, len(cdf)-2] with (cdf[s] & 0xFF) <= q
    (create_matrices_WarpPerRow.py:189-208)."""
    c8 = (np.asarray(cdf).astype(np.int64) & 0xFF)[:-1]
    q = np.arange(256)[:, None]
    ok = c8[None, :] <= q
    idx = np.where(ok, np.arange(len(c8))[None, :], -1).max(axis=1)
    ppf = idx.astype(np.uint8)
    assert ppf[-1] == len(cdf) - 2 and ppf[0] == 0
    return ppf


# -------------------------------------------------------------------- rANS coder
class AnsCoder:
    """rANS coder with 32-bit head and 16-bit words (create_matrices_WarpPerRow.py:211-272)."""

    def __init__(self, precision=PRECISION, word_size=WORD_BITS):
        self.precision = precision
        self.word_size = word_size
        self.bulk = []
        # AnsCoder(precision, 16, compressed=[0, 1]) pops 1 then 0: head = 1 << 16 (:226-228)
        self.head = 1 << word_size

    def push(self, symbol, cdf):
        """:230-248: flush one word if (head >> 24) >= prob, then the bits-back step."""
        prob = int(cdf[symbol + 1]) - int(cdf[symbol])
        if (self.head >> (2 * self.word_size - self.precision)) >= prob:
            while (self.head >> self.word_size) != 0:
                self.bulk.append(self.head & 0xFFFF)
                self.head >>= self.word_size
        z = self.head % prob + int(cdf[symbol])
        self.head = (((self.head // prob) << self.precision) | z) & 0xFFFFFFFF

    def interrupt(self):
        """:267-272: emit low word then high word of the head."""
        for _ in range(2):
            self.bulk.append(self.head & 0xFFFF)
            self.head >>= self.word_size
        self.head = 1 << self.word_size


def lane_columns(cols: int, lane: int):
    """Columns of lane ``lane`` in decode order: k -> 512*(k//16) + 16*lane + k%16
    (create_matrices_WarpPerRow.py:361-366; matvansWarpPerRow-kernel.cu:113-117,133)."""
    k = np.arange(cols // LANES)
    return GROUP_COLS * (k // 16) + 16 * lane + (k % 16)


def encode_matrix(matrix, min_value: int, cdf):
    """Row-interleaved 32-stream encode (create_matrices_WarpPerRow.py:353-430),
    generalised from square to rows x cols (cols % 512 == 0; the reference uses the
    square size for both).  Unused (padding) slots are 0 here; the reference leaves
    ``np.empty`` garbage (:412).  Returns (cursors u32[rows], payload u16, used bool)."""
    matrix = np.asarray(matrix, dtype=np.int8)
    rows, cols = matrix.shape
    assert cols % GROUP_COLS == 0
    streams = [[None] * LANES for _ in range(rows)]
    for lane in range(LANES):
        cidx = lane_columns(cols, lane)
        for r in range(rows):
            coder = AnsCoder()
            for v in matrix[r, cidx][::-1].tolist():
                coder.push(v - min_value, cdf)
            coder.interrupt()
            streams[r][lane] = coder.bulk[::-1]
    rows_words = []
    rows_used = []
    for r in range(rows):
        mx = max(len(s) for s in streams[r])
        k = (mx + GRANULE_WORDS - 1) // GRANULE_WORDS
        row = np.zeros(k * GRANULE_WORDS * LANES, dtype=np.uint16)
        used = np.zeros(row.shape, dtype=bool)
        for lane, s in enumerate(streams[r]):
            j = np.arange(len(s))
            pos = (j // 8) * 256 + lane * 8 + (j % 8)
            row[pos] = s
            used[pos] = True
        rows_words.append(row)
        rows_used.append(used)
    lens = [len(x) for x in rows_words]
    cursors = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint32)
    return cursors, np.concatenate(rows_words), np.concatenate(rows_used)


def decode_matrix(rows, cols, cursors, payload, min_value, cdf_u8, ppf, payload_size=None):
    """Decode every row of a v0 record into an int8 matrix, vectorised over the
    rows*32 streams.  Semantics of matvansWarpPerRow-kernel.cu:95-174 (word order of the
    per-lane FIFO: lane l's word j at cursor + 256*(j//8) + 8*l + j%8; head = w0<<16|w1;
    q = head & 0xFF; s = ppf[q]; w = (int8)(min + s); f = (cdf[s+1]-cdf[s]) & 0xFF;
    head = (head>>8)*f + (q - cdf[s]) mod 2^32; if head < 2^16: head = head<<16 | word),
    which is AnsCoder.pop (create_matrices_WarpPerRow.py:250-264) on the same stream.
    Raises IndexError if a stream would read outside the payload."""
    cursors = np.asarray(cursors, dtype=np.int64)
    payload = np.asarray(payload, dtype=np.uint16)
    cdf = np.asarray(cdf_u8, dtype=np.int64)
    cdf = np.concatenate([cdf, np.zeros(max(0, 257 - len(cdf)), np.int64)])
    ppf = np.asarray(ppf, dtype=np.int64)
    n = rows * LANES
    lane = np.tile(np.arange(LANES, dtype=np.int64), rows)
    base = np.repeat(cursors[:rows], LANES) + lane * 8
    size = len(payload) if payload_size is None else int(payload_size)

    def word(j):
        pos = base + 256 * (j // 8) + (j % 8)
        if pos.size and (pos.max() >= size or pos.min() < 0):
            raise IndexError("stream reads past the payload")
        return payload[pos].astype(np.int64)

    head = (word(np.zeros(n, np.int64)) << 16) | word(np.ones(n, np.int64))
    ptr = np.full(n, 2, dtype=np.int64)
    out = np.empty((rows, cols), dtype=np.int8)
    mn = int(min_value)
    for k in range(cols // LANES):
        q = head & 0xFF
        s = ppf[q]
        col = GROUP_COLS * (k // 16) + 16 * lane + (k % 16)
        out[np.arange(n) // LANES, col] = ((mn + s) & 0xFF).astype(np.uint8).view(np.int8)
        f = (cdf[s + 1] - cdf[s]) & 0xFF
        head = ((head >> 8) * f + (q - cdf[s])) & 0xFFFFFFFF
        need = head < (1 << 16)
        if need.any():
            idx = np.nonzero(need)[0]
            pos = base[idx] + 256 * (ptr[idx] // 8) + (ptr[idx] % 8)
            if pos.max() >= size:
                raise IndexError("stream reads past the payload")
            head[idx] = ((head[idx] << 16) | payload[pos].astype(np.int64)) & 0xFFFFFFFF
            ptr[idx] += 1
    return out


def matvec_i32(matrix, vec):
    """Brute-force integer matvec, exact in int64 then narrowed (|acc| < 2^31 for int8)."""
    return (np.asarray(matrix, np.int64) @ np.asarray(vec, np.int64)).astype(np.int32)


# ------------------------------------------------------------------------ formats
class Record:
    """One compressed matrix of a ``.bin`` v0 file (create_matrices_WarpPerRow.py:274-350)."""

    def __init__(self, rows, cols, grid_spacing, cursors, payload_size, min_value, G, cdf, ppf,
                 payload):
        self.rows, self.cols, self.grid_spacing = rows, cols, grid_spacing
        self.cursors = np.asarray(cursors, np.uint32)
        self.payload_size = payload_size
        self.min_value, self.G = min_value, G
        self.cdf = np.asarray(cdf, np.uint8)
        self.ppf = np.asarray(ppf, np.uint8)
        self.payload = np.asarray(payload, np.uint16)

    def decode(self):
        return decode_matrix(self.rows, self.cols, self.cursors, self.payload, self.min_value,
                             self.cdf, self.ppf, self.payload_size)


def _read_exact(f, n):
    b = f.read(n)
    if len(b) != n:
        raise EOFError("truncated .bin file")
    return b


def read_bin(path):
    """``.bin`` v0 reader (create_matrices_WarpPerRow.py:320-350,440-453;
    matvansWarpPerRow.cu:100-145,225-258).  Skips the vector pad (the C++ reader does not).
    Returns (header dict, vector int8, [Record])."""
    with open(path, "rb") as f:
        num, result_hash, max_wc = struct.unpack("<LLL", _read_exact(f, 12))
        (len_v,) = struct.unpack("<L", _read_exact(f, 4))
        vec = np.frombuffer(_read_exact(f, len_v), dtype=np.int8).copy()
        _read_exact(f, 3 - (len_v + 3) % 4)
        recs = []
        for _ in range(num):
            rows, cols, gs = struct.unpack("<LLf", _read_exact(f, 12))
            cursors = np.frombuffer(_read_exact(f, 4 * rows), dtype="<u4").copy()
            psize, mn, G = struct.unpack("<LbB", _read_exact(f, 6))
            cdf = np.frombuffer(_read_exact(f, G + 1), dtype=np.uint8).copy()
            if (G + 1) % 2 == 1:
                _read_exact(f, 1)
            ppf = np.frombuffer(_read_exact(f, 256), dtype=np.uint8).copy()
            payload = np.frombuffer(_read_exact(f, 2 * psize), dtype="<u2").copy()
            if psize % 2 == 1:
                _read_exact(f, 2)
            recs.append(Record(rows, cols, gs, cursors, psize, mn, G, cdf, ppf, payload))
    hdr = {"num_matrices": num, "result_hash": result_hash, "max_word_count": max_wc}
    return hdr, vec, recs


def write_bin(path, result_hash, vector, records, max_word_count=None):
    """``.bin`` v0 writer (create_matrices_WarpPerRow.py:286-318,440-461)."""
    if max_word_count is None:
        max_word_count = max(
            (4 + r.rows) * 2 + (3 + len(r.cdf)) // 2 + len(r.payload) + len(r.payload) % 2
            for r in records)
    with open(path, "wb") as f:
        f.write(struct.pack("<LLL", len(records), result_hash, max_word_count))
        vector = np.asarray(vector, np.int8)
        f.write(struct.pack("<L", len(vector)))
        f.write(vector.tobytes())
        f.write(b"\0" * (3 - (len(vector) + 3) % 4))
        for r in records:
            f.write(struct.pack("<LLf", r.rows, r.cols, r.grid_spacing))
            f.write(np.asarray(r.cursors, "<u4").tobytes())
            f.write(struct.pack("<LbB", len(r.payload), r.min_value, len(r.cdf) - 1))
            f.write(np.asarray(r.cdf, np.uint8).tobytes())
            if len(r.cdf) % 2 == 1:
                f.write(b"\0")
            f.write(np.asarray(r.ppf, np.uint8).tobytes())
            f.write(np.asarray(r.payload, "<u2").tobytes())
            if len(r.payload) % 2 == 1:
                f.write(b"\0\0")


def read_raw_bin(path):
    """Raw baseline file (create_matrices_WarpPerRow.py:94-129; reader
    baselines/baseline_int8/baseline_matVec_custom.cu:371-393,480-508).
    Returns (num, result_hash, vector int8, [int8 matrices])."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rb") as f:
        num, result_hash = struct.unpack("<LL", _read_exact(f, 8))
        (len_v,) = struct.unpack("<L", _read_exact(f, 4))
        vec = np.frombuffer(_read_exact(f, len_v), dtype=np.int8).copy()
        _read_exact(f, 3 - (len_v + 3) % 4)
        mats = []
        for _ in range(num):
            rows, cols, _gs = struct.unpack("<LLf", _read_exact(f, 12))
            m = np.frombuffer(_read_exact(f, rows * cols), dtype=np.int8).reshape(rows, cols)
            if (rows * cols) % 2 == 1:
                _read_exact(f, 1)
            mats.append(m.copy())
    return num, result_hash, vec, mats


def encode_record(matrix, model=None):
    """Model + ppf + encode for one matrix -> Record (create_matrices_WarpPerRow.py:353-430)."""
    matrix = np.asarray(matrix, np.int8)
    if model is None:
        mn, _G, cdf, _e, _x = create_entropy_model(matrix)
    else:
        mn, cdf = model
    ppf = create_ppf(cdf)
    cursors, payload, _used = encode_matrix(matrix, mn, cdf)
    cdf8 = (np.asarray(cdf).astype(np.int64) & 0xFF).astype(np.uint8)
    return Record(matrix.shape[0], matrix.shape[1], 1.0, cursors, len(payload), mn,
                  len(cdf8) - 1, cdf8, ppf, payload)


# ------------------------------------------------------------- C oracle (ctypes)
_CLIB = None


def c_lib():
    """Load ``oracle/liboracle.so`` (built by ``make -C oracle``)."""
    global _CLIB
    if _CLIB is None:
        import ctypes
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        u32, u64, i8, u8, i32 = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int8,
                                 ctypes.c_uint8, ctypes.c_int)
        lib.oracle_fxhash.argtypes = [P, u64]
        lib.oracle_fxhash.restype = u32
        lib.oracle_decode_v0.argtypes = [u32, u32, P, P, u64, i8, u8, P, P, P]
        lib.oracle_decode_v0.restype = i32
        lib.oracle_matvec_v0.argtypes = [u32, u32, P, P, u64, i8, u8, P, P, P, P, i32]
        lib.oracle_matvec_v0.restype = i32
        lib.oracle_encode_v0.argtypes = [u32, u32, P, i8, u32, P, P, P, u64]
        lib.oracle_encode_v0.restype = ctypes.c_int64
        _CLIB = lib
    return _CLIB


def _p(a):
    return a.ctypes.data


def c_fxhash(arr) -> int:
    a = np.ascontiguousarray(arr, dtype=np.int8)
    return int(c_lib().oracle_fxhash(_p(a), a.size))


def _rec_args(rec):
    cur = np.ascontiguousarray(rec.cursors, np.uint32)
    pay = np.ascontiguousarray(rec.payload, np.uint16)
    cdf = np.zeros(257, np.uint8)
    cdf[:len(rec.cdf)] = rec.cdf
    ppf = np.ascontiguousarray(rec.ppf, np.uint8)
    return cur, pay, cdf, ppf


def c_decode(rec):
    cur, pay, cdf, ppf = _rec_args(rec)
    out = np.empty((rec.rows, rec.cols), np.int8)
    st = c_lib().oracle_decode_v0(rec.rows, rec.cols, _p(cur), _p(pay), int(rec.payload_size),
                                  rec.min_value, rec.G, _p(cdf), _p(ppf), _p(out))
    if st != 0:
        raise ValueError(f"oracle_decode_v0 status {st}")
    return out


def c_matvec(rec, vec, nthreads=1):
    cur, pay, cdf, ppf = _rec_args(rec)
    v = np.ascontiguousarray(vec, np.int8)
    dst = np.empty(rec.rows, np.int32)
    st = c_lib().oracle_matvec_v0(rec.rows, rec.cols, _p(cur), _p(pay), int(rec.payload_size),
                                  rec.min_value, rec.G, _p(cdf), _p(ppf), _p(v), _p(dst),
                                  int(nthreads))
    if st != 0:
        raise ValueError(f"oracle_matvec_v0 status {st}")
    return dst


def c_encode(matrix, min_value, cdf16):
    """C restatement of ``encode_matrix`` -> (cursors, payload)."""
    m = np.ascontiguousarray(matrix, np.int8)
    rows, cols = m.shape
    c16 = np.ascontiguousarray(cdf16, np.uint16)
    cap = rows * 256 * ((cols // 32 + 2 + 7) // 8)
    cur = np.empty(rows, np.uint32)
    pay = np.empty(cap, np.uint16)
    n = c_lib().oracle_encode_v0(rows, cols, _p(m), int(min_value), len(c16) - 1, _p(c16),
                                 _p(cur), _p(pay), cap)
    if n < 0:
        raise ValueError(f"oracle_encode_v0 status {n}")
    return cur, pay[:n].copy()
