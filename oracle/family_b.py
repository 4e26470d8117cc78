"""CPU oracle for the reference's family-B codec ("buffered": quad rANS, 7-bit
probabilities, 64-bit state, 32-bit words) and its two layouts, ans-warp-rows and
ans-warp-stripes.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check results.

Restates buffered/encoder.py and the decoders of buffered/ans-warp-rows-kernel.cu and
buffered/ans-warp-stripes-kernel.cu; each function cites the lines it follows.
Pinned by (``tests/test_oracle_b.py``) the golden safetensors files written by
encoder.py (``tests/golden/make_fixtures.py``):
  gen_matrices ........ the raw matrices, start vector, scales and hashes (byte-exact)
  fit_model ........... the files' ppf / cum / prob tables (byte-exact)
  encode_rows/stripes . the files' compressed words and warp offsets (word-exact)
  decode_rows/stripes . decode(file) == raw matrices; round trips on other matrices
  quad_pop ............ a textbook pop loop (four single-symbol rANS steps) on one coder
  chain ............... the files' per-matrix hashes through the int8 epilogue
"""
from __future__ import annotations

import numpy as np

PRECISION = 7           # DetachedAnsQuadCoder(7, 32, 64) (encoder.py:331,405)
WORD_BITS = 32
STATE_BITS = 64
WARP = 32               # warp_size of the files (encoder.py:441)
DEFAULT_SEED = 20250829  # encoder.py:536


def fxhash(arr) -> int:
    """encoder.py:8-20 (same FxHash as the WarpPerRow path)."""
    h = 0
    for b in np.asarray(arr, np.int8).astype(np.uint8).ravel().tolist():
        h = ((h >> 27) | (h << 5)) & 0xFFFFFFFF
        h = ((h ^ b) * 0x27220A95) & 0xFFFFFFFF
    return h


def gen_matrices(dim, count, target_entropy=4.0, seed=DEFAULT_SEED):
    """create_rnd_matrices_and_vec (encoder.py:23-66) with np.random.seed(seed)
    (:554-555).  Returns (matrices int8[count, dim, dim], initial_vector, scales f32,
    hashes u32)."""
    rs = np.random.RandomState(seed)
    std = 2 ** target_entropy / np.sqrt(2 * np.pi * np.e)
    v0 = np.round(rs.normal(0, 4.0, size=(dim,))).clip(-127.0, 127.0).astype(np.int8)
    mats = np.empty((count, dim, dim), np.int8)
    scales = np.empty((count,), np.float32)
    hashes = np.empty((count,), np.uint32)
    vector = v0
    for i in range(count):
        mats[i] = np.round(rs.normal(0, std, size=(dim, dim))).clip(-127.0, 127.0).astype(np.int8)
        vector = mats[i] @ vector.astype(np.int32)
        scales[i] = 127.0 / np.abs(vector).max()
        vector = np.round(vector * scales[i]).astype(np.int8)   # float64 under numpy 2
        hashes[i] = fxhash(vector)
    return mats, v0, scales, hashes


def fit_model(data):
    """fit_128bit_encoder_model (encoder.py:268-306): counts -> max(1, round_f32(count *
    128 / n)), then the excess is removed one unit at a time from the first maximum
    (np.argmax), or the deficit added to it; cdf uint8 over [min, max] with zero-width
    gaps for absent values.  Returns (min_value, cdf uint8[max-min+2])."""
    data = np.asarray(data, np.int8)
    numel = data.size
    values, counts = np.unique(data, return_counts=True)
    counts = counts.astype(np.float32)
    pmf = np.maximum(1, np.round(counts * 128 / numel).astype(np.uint32))
    total = int(pmf.sum())
    if total > 128:
        for _ in range(total - 128):
            pmf[np.argmax(pmf)] -= 1
    else:
        for _ in range(128 - total):
            pmf[np.argmax(pmf)] += 1
    mn, mx = int(values[0]), int(values[-1])
    cdf = np.zeros((mx - mn + 2,), np.uint8)
    left, last = 0, 0
    for v, f in zip(values.tolist(), pmf.tolist()):
        idx = v - mn
        cdf[last + 1:idx + 1] = left
        last = idx
        left += f
    assert left == 128
    cdf[-1] = left
    return mn, cdf


def decoder_tables(min_value, cdf):
    """decoder_model_from_128bit_encoder_model (encoder.py:309-320): ppf (int8 symbol),
    ppf_cum and ppf_prob over the 128 quantiles."""
    ppf = np.empty(128, np.int8)
    cum = np.empty(128, np.uint8)
    prob = np.empty(128, np.uint8)
    for i, (lo, hi) in enumerate(zip(cdf[:-1].tolist(), cdf[1:].tolist())):
        ppf[lo:hi] = i + min_value
        cum[lo:hi] = lo
        prob[lo:hi] = hi - lo
    return ppf, cum, prob


def pack_4bytes(arr):
    """encoder.py:465-471: little-endian packing, 4 bytes per uint32."""
    a = np.asarray(arr).astype(np.uint8).reshape(-1, 4).astype(np.uint32)
    return a[:, 0] | (a[:, 1] << 8) | (a[:, 2] << 16) | (a[:, 3] << 24)


class QuadCoder:
    """DetachedAnsQuadCoder(7, 32, 64) encode side (encoder.py:113-246)."""

    def __init__(self):
        self.state = 0

    def encode(self, quad, min_value, cdf, buf):
        """encoder.py:198-237: flush a word if four symbols would overflow the state;
        initialise an empty state to total_prob << 4; then the four bits-back steps."""
        s = [int(v) - min_value for v in quad]
        left = [int(cdf[x]) for x in s]
        probs = [int(cdf[x + 1]) - l for x, l in zip(s, left)]
        total = probs[0] * probs[1] * probs[2] * probs[3]
        if (self.state >> (STATE_BITS - 4 * PRECISION)) >= total:
            buf.append(self.state & 0xFFFFFFFF)
            self.state >>= WORD_BITS
        elif self.state == 0:
            self.state = total << (STATE_BITS - WORD_BITS - 4 * PRECISION)
        concat = 0
        for i in range(4):
            rem = self.state % probs[i]
            self.state //= probs[i]
            concat |= (left[i] + rem) << (i * PRECISION)
        self.state = (self.state << (4 * PRECISION)) | concat


def _arrange(bulk_words, coders, W):
    """The per-row / per-stripe rearrangement of encoder.py:346-378 (and :402-425):
    reverse the joint bulk, pad it to 2W mod 4W words, insert the W low then W high
    state words after the first 2W bulk words, and transpose chunks of 4 x W words so
    that each lane reads one uint4 per chunk."""
    bulk = np.asarray(bulk_words[::-1], np.uint32)
    pad = (2 * W - bulk.size) % (4 * W)
    bulk = np.concatenate([bulk, np.zeros(pad, np.uint32)])
    st = np.array([c.state & 0xFFFFFFFF for c in coders] + [c.state >> 32 for c in coders],
                  np.uint32)
    un = np.concatenate([bulk[:2 * W], st, bulk[2 * W:]])
    return un.reshape(-1, 4, W).transpose(0, 2, 1).ravel()


def encode_rows(matrix, W=WARP, model=None):
    """compress_ans_warp_rows (encoder.py:381-430).  The W coders keep their state from
    row to row, as the reference does.  Returns (compressed u32, warp_offsets u32 in
    uint4 units, min_value, cdf)."""
    m = np.asarray(matrix, np.int8)
    out_dim, in_dim = m.shape
    assert in_dim % (4 * W) == 0
    mn, cdf = fit_model(m) if model is None else model
    coders = [QuadCoder() for _ in range(W)]
    offsets = np.empty(out_dim, np.uint32)
    parts, size = [], 0
    rs = m.reshape(out_dim, in_dim // (4 * W), W, 4)
    for i in range(out_dim):
        offsets[i] = size // 4
        buf = []
        for chunk in rs[i][::-1]:
            for coder, quad in zip(coders[::-1], chunk[::-1]):
                coder.encode(quad, mn, cdf, buf)
        row = _arrange(buf, coders, W)
        parts.append(row)
        size += row.size
    return np.concatenate(parts), offsets, mn, cdf


def encode_stripes(matrix, W=WARP, model=None):
    """compress_ans_warp_stripes (encoder.py:323-378): W consecutive rows share a word
    stream; each row is one coder."""
    m = np.asarray(matrix, np.int8)
    out_dim, in_dim = m.shape
    assert out_dim % W == 0 and in_dim % 4 == 0
    mn, cdf = fit_model(m) if model is None else model
    coders = [QuadCoder() for _ in range(W)]
    offsets = np.empty(out_dim // W, np.uint32)
    parts, size = [], 0
    ro = m.reshape(out_dim // W, W, in_dim // 4, 4).transpose(0, 2, 1, 3)
    for i, stripe in enumerate(ro):
        offsets[i] = size // 4
        buf = []
        for quad_col in stripe[::-1]:
            for coder, quad in zip(coders[::-1], quad_col[::-1]):
                coder.encode(quad, mn, cdf, buf)
        part = _arrange(buf, coders, W)
        parts.append(part)
        size += part.size
    return np.concatenate(parts), offsets, mn, cdf


# ---------------------------------------------------------------------- decoders
def _stream_word(compressed, off4, W, pos):
    """Word number ``pos`` of a warp's stream (state words skipped): positions 0..2W-1
    are components x, y of chunk 0; then chunk c >= 1 holds 4W positions in component
    order x (lanes 0..W-1), y, z, w (ans-warp-rows-kernel.cu:62-139)."""
    pos = np.asarray(pos, np.int64)
    c = np.where(pos < 2 * W, 0, 1 + (pos - 2 * W) // (4 * W))
    r = np.where(pos < 2 * W, pos, (pos - 2 * W) % (4 * W))
    comp, lane = r // W, r % W
    return compressed[4 * (off4 + c * W + lane) + comp].astype(np.uint64)


def _decode_warps(compressed, offsets, tables, n_iter, W):
    """Run len(offsets) warps for n_iter quad steps each.  Returns symbols
    int8[n_warps, n_iter, W, 4] (warp, step, lane, symbol of the quad)."""
    ppf, cum, prob = (np.asarray(t, np.int64) for t in tables)
    comp = np.asarray(compressed, np.uint32)
    off = np.asarray(offsets, np.int64)
    n = len(off)
    lanes = np.arange(W)
    lo = comp[4 * (off[:, None] + lanes[None, :]) + 2].astype(np.uint64)
    hi = comp[4 * (off[:, None] + lanes[None, :]) + 3].astype(np.uint64)
    state = (hi << np.uint64(32)) | lo
    cursor = np.zeros(n, np.int64)
    out = np.empty((n, n_iter, W, 4), np.int8)
    for j in range(n_iter):
        need = (state >> np.uint64(32)) == 0
        if need.any():
            pre = np.cumsum(need, axis=1) - need
            pos = cursor[:, None] + pre
            words = _stream_word(comp, off[:, None], W, np.where(need, pos, 0))
            state = np.where(need, (state << np.uint64(32)) | words, state)
            cursor += need.sum(axis=1)
        q = [((state >> np.uint64(7 * i)) & np.uint64(0x7F)).astype(np.int64) for i in range(4)]
        for i in range(4):
            out[:, j, :, i] = ppf[q[i]]
        x = state >> np.uint64(28)
        for i in (3, 2, 1, 0):
            x = x * prob[q[i]].astype(np.uint64) + (q[i] - cum[q[i]]).astype(np.uint64)
        state = x
    return out


def decode_rows(compressed, warp_offsets, tables, in_dim, out_dim, W=WARP):
    """mat_vec_ans_warp_per_row's decode (ans-warp-rows-kernel.cu:62-210): warp = row,
    lane l's quad at step j covers columns 4*(j*W + l) .. +3."""
    s = _decode_warps(compressed, warp_offsets[:out_dim], tables, in_dim // (4 * W), W)
    return s.reshape(out_dim, in_dim)


def decode_stripes(compressed, warp_offsets, tables, in_dim, out_dim, W=WARP):
    """mat_vec_ans_thread_per_row's decode (ans-warp-stripes-kernel.cu:59-214): warp =
    W consecutive rows, lane = row, step j = columns 4j .. 4j+3."""
    s = _decode_warps(compressed, warp_offsets[:out_dim // W], tables, in_dim // 4, W)
    return s.transpose(0, 2, 1, 3).reshape(out_dim, in_dim)


def quad_pop(words, n_quads, cdf):
    """Textbook single-symbol rANS pops for one coder: each quad is four pops of 7 bits
    from the low end, the state being x = q0 + 128*(q1 + 128*(q2 + 128*(q3 + 128*rest)));
    refill with a 32-bit word when the state is below 2^32.  ``words`` = [lo, hi of the
    initial state, then the coder's words in decode order].  Returns symbol indices."""
    cdf = [int(c) for c in cdf]
    x = (words[1] << 32) | words[0]
    p = 2
    out = []
    for _ in range(n_quads):
        if x >> 32 == 0:
            x = (x << 32) | words[p]
            p += 1
        syms, ranks = [], []
        for i in range(4):
            q = (x >> (7 * i)) & 0x7F
            s = max(k for k in range(len(cdf) - 1) if cdf[k] <= q < cdf[k + 1])
            syms.append(s)
            ranks.append((s, q - cdf[s]))
        y = x >> 28
        for s, r in reversed(ranks):
            y = y * (cdf[s + 1] - cdf[s]) + r
        x = y
        out += syms
    return out


def requant_b(acc, scale):
    """Fused epilogue of both kernels (ans-warp-rows-kernel.cu:212-220):
    clamp(rint(float32(acc) * scale), -128, 127)."""
    y = np.asarray(acc, np.int64).astype(np.float32) * np.float32(scale)
    return np.clip(np.rint(y), -128, 127).astype(np.int8)
