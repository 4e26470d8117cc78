"""CPU oracle for the MVX1 format extensions: S interleaved streams per row (32, 64, 128),
P-bit probabilities (8 or 12) and per-row frequency tables.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check results.

The extensions have no reference implementation; they generalise the reference's
WarpPerRow codec (create_matrices_WarpPerRow.py:131-272,353-430 and
matvansWarpPerRow-kernel.cu:95-174), which is the special case S = 32, P = 8, one table
per matrix.  Pinned by (``tests/test_oracle_ext.py``):
  * the special case: ``encode_rows(S=32, P=8, cdf)`` reproduces ``oracle.encode_matrix``
    (itself pinned to the reference's golden payloads) word for word, and
    ``decode_rows`` decodes the golden files;
  * round trips: decode(encode(W)) == W for every (S, P, table mode), on Gaussian,
    Laplace, near-constant and wide-alphabet matrices, constant rows included;
  * closed forms: the model sums to 2^P with every present symbol >= 1; a textbook rANS
    decode of one stream (independent of the vectorised decoder); brute-force matvec;
  * the container: write(read(x)) is byte-identical, and ``decode_rows`` of a record
    written by the C++ library equals its input.

Codec, generalised (32-bit state x, 16-bit words, P-bit quantiles):
  encode push(s): if (x >> (32 - P)) >= f: emit x & 0xFFFF; x >>= 16
                  x = ((x // f) << P) | (x % f + c[s])
  decode:         q = x & (2^P - 1); s = ppf[q]; x = (x >> P) * f[s] + q - c[s]
                  if x < 2^16: x = (x << 16) | next word
Stream l of a row decodes columns 16*S*(k // 16) + 16*l + k % 16 (k = 0 .. cols/S - 1).
A row's payload is K granules of 512 bytes; in each granule every stream owns
GW = 256 / S consecutive words (8 for S = 32, as in the reference layout; 4 for S = 64;
2 for S = 128),
so stream l's word j sits at u16 index cursor + 256*(j // GW) + GW*l + j % GW.

Quad coder (coder = 1; P = 8, one table or one per row): 64-bit state x, 32-bit words.  A stream's
symbols go in quads (decode order); a quad is the four bytes of the low word of x:
  decode: if x < 2^32: x = (x << 32) | next word
          q_i = byte i of x; s_i = ppf[q_i]; F = f0*f1*f2*f3; R = r0 + f0*(r1 + f1*(r2 + f2*r3))
          with r_i = q_i - c[s_i]; x = (x >> 32) * F + R
  encode (reverse quad order, from x = 2^32): if (x >> 32) >= F: emit x & 0xFFFFFFFF,
          x >>= 32; then for i = 0..3: r_i = x % f_i, x //= f_i; x = (x << 32) | sum (c_i+r_i) << 8i
It is the family-B quad coder of buffered/encoder.py:113-246 with 8-bit quantiles (so
x >> 32 is the high word and the quantiles are whole bytes) and one stream per coder
instead of a stream shared by a warp.  Stream words are u32, stored as two u16 of the
payload; word j of stream l is at u16 index cursor + 256*(j // G4) + 2*(G4*l + j % G4)
with G4 = 128 / S words per stream per 512-byte granule; words 0 and 1 are the low and
high words of the initial state.

Quad coder, body and tails (coder = 2; P = 8, one table, S = 32, 64 or 128): the quad
coder's streams, word for word as coder 1, without per-stream padding.  The payload
starts with the row tables, S + 2 u16 per row (zero-padded to 8 u16 in all): K (u32, as
two u16) and the S tail end offsets e[l].  Row r at u16 index cursor (a multiple of 8)
holds the body, the first 4 K words of every stream in K 16-byte-per-stream granules
(word 4 g + i of stream l at u32 cursor/2 + 4 (g S + l) + i; K = the shortest stream's
whole 4-word chunks), then the tails, words 4 K .. of every stream back to back (stream
l's at u32 cursor/2 + 4 K S + e[l-1], e[-1] = 0), zero words to a multiple of four.
Coder 1 pads every stream to the longest of its row and to its granule slot (~0.3 bits
per entry at 4096 x 14336, ~1 at 14336 x 4096); here a stream costs 16 bits of offset.

Outlier split (quad coder, one table): entries with |w| > band are coded as 0 in the
streams and listed per row as records col | (uint8)w << 24 (out_start[rows + 1] offsets,
row-major order); W = decode(streams) + scatter(records).  Pinned by the round trip on
Laplace and Gaussian matrices, the library's split equal to ``split_outliers`` below, and
the forward / backward GPU results against the integer product / dense FP64 of W.

MVX1 container (little-endian):
  "MVX1", u32 version (1; 2 when a matrix has outliers: each record then ends with
  u32 n_outliers and, when nonzero, u32 out_start[rows + 1], u32 out_rec[n_outliers]),
  u32 num_matrices, u32 result_hash, u32 len_v,
  i8 vector[len_v], zero pad to 4 bytes; then per matrix:
  u32 rows, u32 cols, f32 grid_spacing, u32 streams, u32 precision, u32 table_mode,
  u32 coder, i32 min_value, u32 G, u32 payload_size, u32 cursors[rows],
  table_mode 0: u16 cdf[G + 1] | table_mode 1: u8 freq[rows][G], zero pad to 4 bytes,
  u16 payload[payload_size], zero pad to 4 bytes.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"MVX1"
VERSION = 1


# ------------------------------------------------------------------------ models
def quantize_counts(counts, total):
    """Quantise symbol counts to integer frequencies summing to ``total``.

    The reference's steps (create_matrices_WarpPerRow.py:145-170) with a deterministic
    order: entries sorted by count descending, equal counts by symbol index descending;
    ``max(1, round_half_even(count * total / n))`` for every present symbol; an excess is
    shaved one unit at a time from the leading entries that are above 1, a deficit is
    spread over the trailing entries.  Absent symbols get 0.  Returns int64[len(counts)]
    or raises ValueError if the counts cannot be normalised."""
    counts = np.asarray(counts, np.int64)
    idx = np.nonzero(counts)[0]
    n = int(counts.sum())
    if n == 0:
        raise ValueError("no symbols")
    order = idx[np.lexsort((-idx, -counts[idx]))]
    q = np.maximum(1, np.round(counts[order] * (total / n)).astype(np.int64))
    excess = int(q.sum()) - total
    while excess > 0:
        nred = int((q > 1).sum())
        nr = min(excess, nred)
        if nr <= 0:
            raise ValueError("model cannot be normalised")
        q[:nr] -= 1
        excess -= nr
    if excess < 0:
        if -excess > len(q):
            raise ValueError("model cannot be normalised")
        q[excess:] += 1
    f = np.zeros(len(counts), np.int64)
    f[order] = q
    assert int(f.sum()) == total
    return f


def global_model(matrix, precision):
    """One table for the matrix: (min_value, G, cdf uint16[G + 1]).  A constant matrix
    gets a dummy second symbol of frequency 1 (as create_matrices_WarpPerRow.py:135-143)."""
    m = np.asarray(matrix, np.int8)
    mn, mx = int(m.min()), int(m.max())
    total = 1 << precision
    if mn == mx:
        return mn, 2, np.array([0, total - 1, total], np.uint16)
    G = mx - mn + 1
    counts = np.bincount((m.astype(np.int64) - mn).ravel(), minlength=G)
    f = quantize_counts(counts, total)
    cdf = np.concatenate([[0], np.cumsum(f)]).astype(np.uint16)
    return mn, G, cdf


def row_freqs(row_symbols, G):
    """Per-row 8-bit frequencies over the matrix alphabet (G >= 2).  A row with a single
    distinct symbol s gets f[s] = 255 and a dummy neighbour (s + 1, or s - 1 at the top)
    with f = 1, so that every frequency fits a byte."""
    counts = np.bincount(np.asarray(row_symbols, np.int64), minlength=G)
    present = np.nonzero(counts)[0]
    if len(present) == 1:
        s = int(present[0])
        f = np.zeros(G, np.int64)
        f[s] = 255
        f[s + 1 if s + 1 < G else s - 1] = 1
        return f.astype(np.uint8)
    return quantize_counts(counts, 256).astype(np.uint8)


def per_row_model(matrix):
    """(min_value, G, freq uint8[rows][G]) with one 8-bit table per row."""
    m = np.asarray(matrix, np.int8)
    mn, mx = int(m.min()), int(m.max())
    G = max(2, mx - mn + 1)
    sym = m.astype(np.int64) - mn
    return mn, G, np.stack([row_freqs(sym[r], G) for r in range(m.shape[0])])


def cdf_of(freq):
    return np.concatenate([[0], np.cumsum(np.asarray(freq, np.int64))])


def ppf_of(cdf, precision):
    """ppf[q] = the symbol s with cdf[s] <= q < cdf[s + 1] (never a zero-frequency one)."""
    cdf = np.asarray(cdf, np.int64)
    return (np.searchsorted(cdf, np.arange(1 << precision), side="right") - 1).astype(np.int64)


def lut_of(cdf, min_value, precision):
    """Fused decode table over quantiles: A[q] = f | (uint8)(min + s) << 24,
    B[q] = q - cdf[s]; returns uint32[2 * 2^P] = A then B."""
    cdf = np.asarray(cdf, np.int64)
    ppf = ppf_of(cdf, precision)
    q = np.arange(1 << precision)
    f = cdf[ppf + 1] - cdf[ppf]
    sym = (min_value + ppf) & 0xFF
    return np.concatenate([f | (sym << 24), q - cdf[ppf]]).astype(np.uint32)


# ------------------------------------------------------------------------- coder
def stream_columns(cols, streams, lane):
    k = np.arange(cols // streams)
    return 16 * streams * (k // 16) + 16 * lane + (k % 16)


def encode_stream(symbols, cdf, precision):
    """Encode one stream (symbols in decode order); returns its words in decode order:
    [x_hi, x_lo, last flushed, ..., first flushed] (AnsCoder push/interrupt,
    create_matrices_WarpPerRow.py:226-272, with P-bit probabilities)."""
    cdf = [int(c) for c in cdf]
    x = 1 << 16
    bulk = []
    for s in reversed([int(v) for v in symbols]):
        f = cdf[s + 1] - cdf[s]
        if (x >> (32 - precision)) >= f:
            bulk.append(x & 0xFFFF)
            x >>= 16
        x = ((x // f) << precision) | (x % f + cdf[s])
    bulk.append(x & 0xFFFF)
    bulk.append(x >> 16)
    return bulk[::-1]


def encode_stream_quad(symbols, cdf):
    """Quad coder (coder = 1) for one stream, symbols in decode order (a multiple of
    four); returns its u32 words in decode order [x_lo, x_hi, last flushed, ...]."""
    cdf = [int(c) for c in cdf]
    sym = [int(v) for v in symbols]
    x = 1 << 32
    bulk = []
    for j in reversed(range(len(sym) // 4)):
        quad = sym[4 * j:4 * j + 4]
        f = [cdf[t + 1] - cdf[t] for t in quad]
        F = f[0] * f[1] * f[2] * f[3]
        if (x >> 32) >= F:
            bulk.append(x & 0xFFFFFFFF)
            x >>= 32
        concat = 0
        for i in range(4):
            concat |= (cdf[quad[i]] + x % f[i]) << (8 * i)
            x //= f[i]
        x = (x << 32) | concat
    return [x & 0xFFFFFFFF, x >> 32] + bulk[::-1]


def encode_rows_quad(matrix, streams, min_value, cdf=None, freq=None):
    """Quad-coder encode of a matrix with one table (``cdf``) or one per row (``freq``).
    Returns (cursors uint32[rows] in u16 units, payload uint16)."""
    m = np.asarray(matrix, np.int8)
    rows, cols = m.shape
    assert streams in (32, 64, 128) and cols % (16 * streams) == 0
    G4 = 128 // streams
    out, cursors, pos = [], [], 0
    for r in range(rows):
        c = cdf if freq is None else cdf_of(freq[r])
        sym = m[r].astype(np.int64) - min_value
        lanes = [encode_stream_quad(sym[stream_columns(cols, streams, l)], c)
                 for l in range(streams)]
        K = (max(len(s) for s in lanes) + G4 - 1) // G4
        row = np.zeros(K * 128, np.uint32)
        for l, s in enumerate(lanes):
            j = np.arange(len(s))
            row[128 * (j // G4) + G4 * l + j % G4] = s
        cursors.append(pos)
        pos += 2 * len(row)
        out.append(row.view(np.uint16))
    return np.asarray(cursors, np.uint32), np.concatenate(out)


def decode_rows_quad(rows, cols, streams, cursors, payload, min_value, cdf=None, freq=None):
    """Decode a quad-coder matrix (one table ``cdf`` or one per row ``freq``), vectorised
    over the rows x streams streams."""
    pay32 = np.asarray(payload, np.uint16).view(np.uint32)
    S, G4 = streams, 128 // streams
    if freq is None:
        cdfs = np.tile(np.asarray(cdf, np.int64), (rows, 1))
    else:
        cdfs = np.stack([cdf_of(freq[r]) for r in range(rows)])
    ppfs = np.stack([ppf_of(c, 8) for c in cdfs])
    n = rows * S
    row = np.repeat(np.arange(rows), S)
    lane = np.tile(np.arange(S), rows)
    base = np.repeat(np.asarray(cursors, np.int64)[:rows] // 2, S) + G4 * lane

    def word(b, j):
        p = b + 128 * (j // G4) + j % G4
        if p.size and p.max() >= len(pay32):
            raise IndexError("stream reads past the payload")
        return pay32[p].astype(np.uint64)

    x = (word(base, np.ones(n, np.int64)) << np.uint64(32)) | word(base, np.zeros(n, np.int64))
    ptr = np.full(n, 2, np.int64)
    out = np.empty((rows, cols), np.int8)
    for j in range(cols // S // 4):
        need = x < (np.uint64(1) << np.uint64(32))
        if need.any():
            i = np.nonzero(need)[0]
            x[i] = (x[i] << np.uint64(32)) | word(base[i], ptr[i])
            ptr[i] += 1
        q = [((x >> np.uint64(8 * i)) & np.uint64(0xFF)).astype(np.int64) for i in range(4)]
        F = np.ones(n, np.uint64)
        R = np.zeros(n, np.uint64)
        mul = np.ones(n, np.uint64)
        for i in range(4):
            s = ppfs[row, q[i]]
            k = 4 * j + i
            col = 16 * S * (k // 16) + 16 * lane + k % 16
            out[row, col] = ((min_value + s) & 0xFF).astype(np.uint8).view(np.int8)
            f = (cdfs[row, s + 1] - cdfs[row, s]).astype(np.uint64)
            R = R + mul * (q[i] - cdfs[row, s]).astype(np.uint64)
            mul = mul * f
        x = (x >> np.uint64(32)) * mul + R
    return out


def encode_rows(matrix, streams, precision, min_value, cdf=None, freq=None):
    """Encode a matrix with ``streams`` streams per row, one table (``cdf``) or one per
    row (``freq``).  Returns (cursors uint32[rows], payload uint16)."""
    m = np.asarray(matrix, np.int8)
    rows, cols = m.shape
    assert streams in (32, 64, 128) and cols % (16 * streams) == 0
    GW = 256 // streams
    out, cursors, pos = [], [], 0
    for r in range(rows):
        c = cdf if freq is None else cdf_of(freq[r])
        sym = m[r].astype(np.int64) - min_value
        lanes = [encode_stream(sym[stream_columns(cols, streams, l)], c, precision)
                 for l in range(streams)]
        K = (max(len(s) for s in lanes) + GW - 1) // GW
        row = np.zeros(K * 256, np.uint16)
        for l, s in enumerate(lanes):
            j = np.arange(len(s))
            row[256 * (j // GW) + GW * l + j % GW] = s
        cursors.append(pos)
        pos += len(row)
        out.append(row)
    return np.asarray(cursors, np.uint32), np.concatenate(out)


def lanes_table_u16(rows, streams):
    """u16 words of a coder-2 matrix's row tables (rows x (S + 2), zero-padded to 8)."""
    return (rows * (streams + 2) + 7) // 8 * 8


def encode_rows_quad_lanes(matrix, streams, min_value, cdf):
    """Quad-coder encode of a matrix in coder 2's layout (one table): per row, the first
    4 K words of every stream in K granules (K = the shortest stream's whole 4-word
    chunks), the rest of each stream back to back.  Returns (cursors uint32[rows] in u16
    units, payload uint16)."""
    m = np.asarray(matrix, np.int8)
    rows, cols = m.shape
    S = streams
    assert S in (32, 64, 128) and cols % (16 * S) == 0
    table = np.zeros(lanes_table_u16(rows, S), np.uint16)
    out, cursors, pos = [], [], len(table)
    for r in range(rows):
        sym = m[r].astype(np.int64) - min_value
        lanes = [encode_stream_quad(sym[stream_columns(cols, S, l)], cdf) for l in range(S)]
        K = min(len(s) for s in lanes) // 4
        body = np.array([s[:4 * K] for s in lanes], np.uint32).reshape(S, K, 4)
        tails = [s[4 * K:] for s in lanes]
        ends = np.cumsum([len(t) for t in tails])
        if ends[-1] > 0xFFFF or K > 0xFFFF:
            raise ValueError("row %d: tails do not fit 16-bit offsets" % r)
        t = table[r * (S + 2):(r + 1) * (S + 2)]
        t[0], t[1], t[2:] = K & 0xFFFF, K >> 16, ends
        words = np.concatenate([body.transpose(1, 0, 2).ravel(),
                                np.asarray(sum(tails, []), np.uint32)])
        row = np.zeros((len(words) + 3) // 4 * 4, np.uint32)
        row[:len(words)] = words
        cursors.append(pos)
        pos += 2 * len(row)
        out.append(row.view(np.uint16))
    return np.asarray(cursors, np.uint32), np.concatenate([table] + out)


def lanes_row(payload, cursors, rows, streams, r):
    """(K, tail starts, tail lengths) of row r of a coder-2 matrix; the starts are u32
    indices into the payload."""
    S = streams
    pay = np.asarray(payload, np.uint16)
    t = pay[r * (S + 2):(r + 1) * (S + 2)].astype(np.int64)
    K = int(t[0]) | int(t[1]) << 16
    ends = t[2:]
    begs = np.concatenate([[0], ends[:-1]])
    if np.any(ends < begs):
        raise ValueError("tail offsets decrease")
    tail0 = int(cursors[r]) // 2 + K * S * 4
    return K, tail0 + begs, ends - begs


def lane_words(payload, cursors, rows, streams, r, l):
    """The u32 words of stream l of row r of a coder-2 matrix, in decode order."""
    pay32 = np.asarray(payload, np.uint16).view(np.uint32)
    K, tb, tn = lanes_row(payload, cursors, rows, streams, r)
    base = int(cursors[r]) // 2
    body = [int(pay32[base + (g * streams + l) * 4 + i]) for g in range(K) for i in range(4)]
    return body + [int(w) for w in pay32[tb[l]:tb[l] + tn[l]]]


def decode_rows_quad_lanes(rows, cols, streams, cursors, payload, min_value, cdf):
    """Decode a coder-2 matrix (one table), vectorised over the rows x streams streams.
    A stream reading past its own words raises."""
    pay32 = np.asarray(payload, np.uint16).view(np.uint32)
    S = streams
    Ks, tb, tn, base = [], [], [], []
    for r in range(rows):
        K, b, n = lanes_row(payload, cursors, rows, S, r)
        Ks.append(np.full(S, K))
        tb.append(b)
        tn.append(n)
        base.append(int(cursors[r]) // 2 + 4 * np.arange(S))
    Ks, tb, tn, base = (np.concatenate(a).astype(np.int64) for a in (Ks, tb, tn, base))
    if tb.size and (tb + tn).max() > len(pay32):   # tails follow the bodies
        raise IndexError("stream past the payload")

    def word(i, j):
        if np.any(j >= 4 * Ks[i] + tn[i]):
            raise IndexError("stream reads past its words")
        body = j < 4 * Ks[i]
        p = np.where(body, base[i] + (j // 4) * 4 * S + j % 4, tb[i] + j - 4 * Ks[i])
        return pay32[p].astype(np.uint64)

    return _decode_quad(rows, cols, S, np.asarray(cdf, np.int64), min_value, word)


def _decode_quad(rows, cols, S, cdf, min_value, word):
    """Quad-coder decode of rows x S streams (one table) with word(i, j) = word j of
    streams i (row-major stream index)."""
    n = rows * S
    ppf = ppf_of(cdf, 8)
    lane = np.tile(np.arange(S), rows)
    row = np.repeat(np.arange(rows), S)
    allst = np.arange(n)
    x = (word(allst, np.ones(n, np.int64)) << np.uint64(32)) | word(allst, np.zeros(n, np.int64))
    ptr = np.full(n, 2, np.int64)
    out = np.empty((rows, cols), np.int8)
    for j in range(cols // S // 4):
        need = x < (np.uint64(1) << np.uint64(32))
        if need.any():
            i = np.nonzero(need)[0]
            x[i] = (x[i] << np.uint64(32)) | word(i, ptr[i])
            ptr[i] += 1
        F = np.ones(n, np.uint64)
        R = np.zeros(n, np.uint64)
        for i in range(4):
            q = ((x >> np.uint64(8 * i)) & np.uint64(0xFF)).astype(np.int64)
            s = ppf[q]
            k = 4 * j + i
            col = 16 * S * (k // 16) + 16 * lane + k % 16
            out[row, col] = ((min_value + s) & 0xFF).astype(np.uint8).view(np.int8)
            R = R + F * (q - cdf[s]).astype(np.uint64)
            F = F * (cdf[s + 1] - cdf[s]).astype(np.uint64)
        x = (x >> np.uint64(32)) * F + R
    return out


def decode_rows(rows, cols, streams, precision, cursors, payload, min_value, cdf=None,
                freq=None, payload_size=None):
    """Decode every row, vectorised over the rows x streams streams.  Raises IndexError
    if a stream would read outside the payload."""
    payload = np.asarray(payload, np.uint16)
    size = len(payload) if payload_size is None else int(payload_size)
    S, P = streams, precision
    mask = (1 << P) - 1
    if freq is None:
        cdfs = np.broadcast_to(np.asarray(cdf, np.int64), (rows, len(cdf)))
    else:
        cdfs = np.stack([cdf_of(freq[r]) for r in range(rows)])
    ppfs = np.stack([ppf_of(cdfs[r], P) for r in range(rows)])
    n = rows * S
    row = np.repeat(np.arange(rows), S)
    lane = np.tile(np.arange(S), rows)
    base = np.repeat(np.asarray(cursors, np.int64)[:rows], S) + (256 // S) * lane

    x = ((_words(payload, size, base, S, np.zeros(n, np.int64)) << 16)
         | _words(payload, size, base, S, np.ones(n, np.int64)))
    ptr = np.full(n, 2, np.int64)
    out = np.empty((rows, cols), np.int8)
    for k in range(cols // S):
        q = x & mask
        s = ppfs[row, q]
        col = 16 * S * (k // 16) + 16 * lane + k % 16
        out[row, col] = ((min_value + s) & 0xFF).astype(np.uint8).view(np.int8)
        c0 = cdfs[row, s]
        f = cdfs[row, s + 1] - c0
        x = (x >> P) * f + q - c0
        need = x < (1 << 16)
        if need.any():
            i = np.nonzero(need)[0]
            x[i] = (x[i] << 16) | _words(payload, size, base[i], S, ptr[i])
            ptr[i] += 1
    return out


def _words(payload, size, base, S, j):
    GW = 256 // S
    p = base + 256 * (j // GW) + j % GW
    if p.size and p.max() >= size:
        raise IndexError("stream reads past the payload")
    return payload[p].astype(np.int64)


def decode_stream_textbook(words, n, cdf, precision):
    """Textbook rANS decode of one stream given as a list of words in decode order
    (state = w0 << 16 | w1, then pop); independent of the layout and the vectorised
    decoder.  Returns the n symbol indices."""
    cdf = [int(c) for c in cdf]
    x = (words[0] << 16) | words[1]
    p = 2
    out = []
    M = 1 << precision
    for _ in range(n):
        q = x % M
        s = max(i for i in range(len(cdf) - 1) if cdf[i] <= q and cdf[i + 1] > q)
        out.append(s)
        x = (cdf[s + 1] - cdf[s]) * (x // M) + q - cdf[s]
        if x < 1 << 16:
            x = x * 65536 + words[p]
            p += 1
    return out


def decode_stream_quad_textbook(words, n_quads, cdf):
    """One quad-coder stream (u32 words in decode order, [x_lo, x_hi, ...]) decoded by
    four plain 8-bit rANS pops per quad, x = q0 + 256*(q1 + 256*(q2 + 256*(q3 + 256*y))):
    pop i takes q_i = x % 256 and leaves y_i = f(s_i) * (x // 256) + q_i - cdf[s_i] --
    applied from the last quantile inwards, which is the quad coder's closed form.
    Independent of the vectorised decoder; returns the symbol indices."""
    cdf = [int(c) for c in cdf]
    x = (words[1] << 32) | words[0]
    p = 2
    out = []
    for _ in range(n_quads):
        if x < 1 << 32:
            x = (x << 32) | words[p]
            p += 1
        qs = [(x >> (8 * i)) & 0xFF for i in range(4)]
        ss = [max(k for k in range(len(cdf) - 1) if cdf[k] <= q < cdf[k + 1]) for q in qs]
        y = x >> 32
        for q, t in zip(reversed(qs), reversed(ss)):
            y = (cdf[t + 1] - cdf[t]) * y + q - cdf[t]
        x = y
        out += ss
    return out


# ---------------------------------------------------------------------- outliers
def split_outliers(matrix, band):
    """(coded matrix: entries with |w| > band replaced by 0, out_start uint32[rows + 1],
    out_rec uint32[n] = col | (uint8)w << 24, rows in order, columns ascending)."""
    m = np.asarray(matrix, np.int8)
    out = np.abs(m.astype(np.int16)) > band
    rows_i, cols_i = np.nonzero(out)           # row-major order
    rec = (cols_i.astype(np.uint32) |
           (m[rows_i, cols_i].view(np.uint8).astype(np.uint32) << 24)).astype(np.uint32)
    start = np.zeros(m.shape[0] + 1, np.uint32)
    np.cumsum(out.sum(axis=1), out=start[1:])
    coded = m.copy()
    coded[out] = 0
    return coded, start, rec


def merge_outliers(coded, out_start, out_rec):
    """The matrix W from the coded matrix and its outlier records."""
    w = np.array(coded, np.int8, copy=True)
    for r in range(w.shape[0]):
        for o in range(int(out_start[r]), int(out_start[r + 1])):
            w[r, int(out_rec[o]) & 0xFFFFFF] = np.int8(np.uint8(int(out_rec[o]) >> 24))
    return w


# --------------------------------------------------------------------- container
class RecordV1:
    def __init__(self, rows, cols, streams, precision, table_mode, min_value, G, cursors,
                 payload, cdf=None, freq=None, grid_spacing=1.0, coder=0, out_start=None,
                 out_rec=None):
        self.coder = coder
        self.out_start = None if out_start is None else np.asarray(out_start, np.uint32)
        self.out_rec = None if out_rec is None else np.asarray(out_rec, np.uint32)
        self.rows, self.cols, self.streams, self.precision = rows, cols, streams, precision
        self.table_mode, self.min_value, self.G = table_mode, min_value, G
        self.cursors = np.asarray(cursors, np.uint32)
        self.payload = np.asarray(payload, np.uint16)
        self.cdf = None if cdf is None else np.asarray(cdf, np.uint16)
        self.freq = None if freq is None else np.asarray(freq, np.uint8).reshape(rows, G)
        self.grid_spacing = grid_spacing

    def decode(self):
        if self.coder == 2:
            return decode_rows_quad_lanes(self.rows, self.cols, self.streams, self.cursors,
                                          self.payload, self.min_value, self.cdf)
        if self.coder == 1:
            return decode_rows_quad(self.rows, self.cols, self.streams, self.cursors,
                                    self.payload, self.min_value, self.cdf, self.freq)
        return decode_rows(self.rows, self.cols, self.streams, self.precision, self.cursors,
                           self.payload, self.min_value, self.cdf, self.freq)


def encode_record(matrix, streams=64, precision=8, table_mode=0, coder=0):
    m = np.asarray(matrix, np.int8)
    if coder == 2:
        assert precision == 8 and table_mode == 0
        mn, G, cdf = global_model(m, 8)
        cur, pay = encode_rows_quad_lanes(m, streams, mn, cdf)
        return RecordV1(m.shape[0], m.shape[1], streams, 8, 0, mn, G, cur, pay, cdf=cdf,
                        coder=2)
    if coder == 1:
        assert precision == 8
        if table_mode == 0:
            mn, G, cdf = global_model(m, 8)
            cur, pay = encode_rows_quad(m, streams, mn, cdf=cdf)
            return RecordV1(m.shape[0], m.shape[1], streams, 8, 0, mn, G, cur, pay, cdf=cdf,
                            coder=1)
        mn, G, freq = per_row_model(m)
        cur, pay = encode_rows_quad(m, streams, mn, freq=freq)
        return RecordV1(m.shape[0], m.shape[1], streams, 8, 1, mn, G, cur, pay, freq=freq,
                        coder=1)
    if table_mode == 0:
        mn, G, cdf = global_model(m, precision)
        cur, pay = encode_rows(m, streams, precision, mn, cdf=cdf)
        return RecordV1(m.shape[0], m.shape[1], streams, precision, 0, mn, G, cur, pay, cdf=cdf)
    assert precision == 8
    mn, G, freq = per_row_model(m)
    cur, pay = encode_rows(m, streams, precision, mn, freq=freq)
    return RecordV1(m.shape[0], m.shape[1], streams, 8, 1, mn, G, cur, pay, freq=freq)


def _pad4(f, n):
    f.write(b"\0" * ((4 - n % 4) % 4))


def write_mvx1(path, records, vector=np.zeros(0, np.int8), result_hash=0):
    with open(path, "wb") as f:
        vector = np.asarray(vector, np.int8)
        f.write(MAGIC + struct.pack("<LLLL", VERSION, len(records), result_hash, len(vector)))
        f.write(vector.tobytes())
        _pad4(f, len(vector))
        for r in records:
            f.write(struct.pack("<LLfLLLLlLL", r.rows, r.cols, r.grid_spacing, r.streams,
                                r.precision, r.table_mode, r.coder, r.min_value, r.G,
                                len(r.payload)))
            f.write(np.asarray(r.cursors, "<u4").tobytes())
            tab = (np.asarray(r.cdf, "<u2").tobytes() if r.table_mode == 0
                   else np.asarray(r.freq, np.uint8).tobytes())
            f.write(tab)
            _pad4(f, len(tab))
            f.write(np.asarray(r.payload, "<u2").tobytes())
            _pad4(f, 2 * len(r.payload))


def read_mvx1(path):
    """Returns (header dict, vector, [RecordV1])."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != MAGIC:
        raise ValueError("not an MVX1 file")
    ver, num, rhash, lv = struct.unpack_from("<LLLL", data, 4)
    if ver not in (1, 2):
        raise ValueError("unsupported MVX1 version")
    o = 20
    vec = np.frombuffer(data, np.int8, lv, o).copy()
    o += lv + (4 - lv % 4) % 4
    recs = []
    for _ in range(num):
        rows, cols, gs, S, P, mode, coder, mn, G, psize = struct.unpack_from("<LLfLLLLlLL", data, o)
        o += 40
        cur = np.frombuffer(data, "<u4", rows, o).copy()
        o += 4 * rows
        if mode == 0:
            cdf = np.frombuffer(data, "<u2", G + 1, o).copy()
            n, freq = 2 * (G + 1), None
        else:
            freq = np.frombuffer(data, np.uint8, rows * G, o).copy().reshape(rows, G)
            n, cdf = rows * G, None
        o += n + (4 - n % 4) % 4
        pay = np.frombuffer(data, "<u2", psize, o).copy()
        o += 2 * psize + (4 - (2 * psize) % 4) % 4
        ost = orec = None
        if ver == 2:
            (n_out,) = struct.unpack_from("<L", data, o)
            o += 4
            if n_out:
                ost = np.frombuffer(data, "<u4", rows + 1, o).copy()
                o += 4 * (rows + 1)
                orec = np.frombuffer(data, "<u4", n_out, o).copy()
                o += 4 * n_out
        recs.append(RecordV1(rows, cols, S, P, mode, mn, G, cur, pay, cdf, freq, gs, coder,
                             ost, orec))
    if o != len(data):
        raise ValueError("trailing bytes")
    return {"num_matrices": num, "result_hash": rhash}, vec, recs
