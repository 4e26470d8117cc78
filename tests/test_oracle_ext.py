"""Pins the MVX1 extension oracle (oracle/ext.py): the reference special case against the
golden-pinned v0 oracle and the golden files, round trips for every (streams, precision,
table mode), a textbook decode of single streams, the models' invariants and the
container's byte round trip."""
import os

import numpy as np
import pytest

from oracle import ext as E
from oracle import oracle as O

from .helpers import gaussian_int8, laplace_int8

VARIANTS = [(32, 8, 0), (64, 8, 0), (128, 8, 0), (32, 12, 0), (64, 12, 0), (128, 12, 0),
            (64, 8, 1)]


def _cases():
    rng = np.random.default_rng(7)
    wide = rng.integers(-127, 128, (5, 2048)).astype(np.int8)
    near = np.zeros((6, 1024), np.int8)
    near[1, 17] = 3
    near[4, :] = -127          # a constant row at the alphabet edge
    near[5, ::7] = 1
    return {
        "gauss6": gaussian_int8(9, 2048, 6.0, seed=1)[0],
        "laplace4": laplace_int8(7, 3072, 4.0, seed=2)[0],
        "gauss0.4": gaussian_int8(4, 1024, 0.4, seed=3)[0],
        "near_const": near,
        "wide": wide,
        "const": np.full((3, 1024), 9, np.int8),
    }


CASES = _cases()


def test_special_case_is_reference_encoder():
    """S = 32, P = 8, one table: word for word the v0 encoder (pinned to the golden
    payloads by tests/test_oracle.py)."""
    for name, m in CASES.items():
        mn, G, cdf, _, _ = O.create_entropy_model(m)
        c0, p0, _ = O.encode_matrix(m, mn, cdf)
        c1, p1 = E.encode_rows(m, 32, 8, mn, cdf=cdf)
        assert np.array_equal(c0, c1) and np.array_equal(p0, p1), name


@pytest.mark.parametrize("fx", ["compressed_matrices_512_5bit_count2.bin",
                                "compressed_matrices_1024_5bit_count1.bin"])
def test_decoder_reads_golden_files(golden_dir, fx):
    hdr, vec, recs = O.read_bin(os.path.join(golden_dir, fx))
    for r in recs:
        cdf = r.cdf.astype(np.int64)
        cdf[-1] = 256
        got = E.decode_rows(r.rows, r.cols, 32, 8, r.cursors, r.payload, r.min_value, cdf=cdf)
        assert np.array_equal(got, O.decode_matrix(r.rows, r.cols, r.cursors, r.payload,
                                                   r.min_value, r.cdf, r.ppf))


def _fits(S, name):
    """The layout needs the column count to be a multiple of 16 * streams."""
    return CASES[name].shape[1] % (16 * S) == 0


@pytest.mark.parametrize("S,P,mode,name", [(S, P, mode, name) for S, P, mode in VARIANTS
                                           for name in CASES if _fits(S, name)])
def test_round_trip(S, P, mode, name):
    m = CASES[name]
    r = E.encode_record(m, S, P, mode)
    assert r.payload.size % 256 == 0 and np.all(r.cursors % 256 == 0)
    assert np.array_equal(r.decode(), m)
    v = np.random.default_rng(3).integers(-127, 128, m.shape[1]).astype(np.int8)
    assert np.array_equal(O.matvec_i32(r.decode(), v),
                          (m.astype(np.int64) @ v.astype(np.int64)).astype(np.int32))


@pytest.mark.parametrize("S,P,mode", VARIANTS)
def test_textbook_decode_of_single_streams(S, P, mode):
    """Gather one stream's words with the layout formula and decode them with a plain
    pop loop: the symbols are the row's columns 16*S*(k//16) + 16*l + k%16."""
    m = CASES["laplace4"][:, :2048] if S >= 64 else CASES["laplace4"]
    r = E.encode_record(m, S, P, mode)
    GW = 256 // S
    for row in (0, m.shape[0] - 1):
        cdf = r.cdf if mode == 0 else E.cdf_of(r.freq[row])
        for lane in (0, 5, S - 1):
            n = m.shape[1] // S
            words = []
            j = 0
            end = r.cursors[row + 1] if row + 1 < m.shape[0] else len(r.payload)
            while True:
                p = int(r.cursors[row]) + 256 * (j // GW) + GW * lane + j % GW
                if p >= end:
                    break
                words.append(int(r.payload[p]))
                j += 1
            syms = E.decode_stream_textbook(words, n, cdf, P)
            want = m[row, E.stream_columns(m.shape[1], S, lane)].astype(np.int64) - r.min_value
            assert syms == want.tolist()


@pytest.mark.parametrize("S,mode,name", [(S, mode, name)
                                          for S, mode in [(32, 0), (64, 0), (128, 0), (64, 1)]
                                          for name in CASES if _fits(S, name)])
def test_quad_round_trip(S, mode, name):
    m = CASES[name]
    r = E.encode_record(m, S, 8, mode, coder=1)
    assert r.payload.size % 256 == 0 and np.all(r.cursors % 256 == 0)
    assert np.array_equal(r.decode(), m)


def test_quad_per_row_tables_textbook():
    """Quad coder with one table per row: each stream decodes with plain 8-bit pops over
    its own row's table (and not with another row's)."""
    m = CASES["laplace4"][:, :2048]
    r = E.encode_record(m, 64, 8, 1, coder=1)
    p32 = r.payload.view(np.uint32)
    for row in range(m.shape[0]):
        end = (r.cursors[row + 1] if row + 1 < m.shape[0] else len(r.payload)) // 2
        words, j = [], 0
        while True:
            p = int(r.cursors[row]) // 2 + 128 * (j // 2) + 2 * 5 + j % 2
            if p >= end:
                break
            words.append(int(p32[p]))
            j += 1
        syms = E.decode_stream_quad_textbook(words, m.shape[1] // 64 // 4, E.cdf_of(r.freq[row]))
        want = m[row, E.stream_columns(m.shape[1], 64, 5)].astype(np.int64) - r.min_value
        assert syms == want.tolist()


@pytest.mark.parametrize("S", [32, 64, 128])
def test_quad_textbook_decode(S):
    """Gather one quad stream's u32 words with the layout formula and decode them with
    plain 8-bit pops."""
    m = CASES["laplace4"][:, :2048] if S >= 64 else CASES["laplace4"]
    r = E.encode_record(m, S, 8, 0, coder=1)
    p32 = r.payload.view(np.uint32)
    G4 = 128 // S
    for row in (0, m.shape[0] - 1):
        end = (r.cursors[row + 1] if row + 1 < m.shape[0] else len(r.payload)) // 2
        for lane in (0, 3, S - 1):
            words, j = [], 0
            while True:
                p = int(r.cursors[row]) // 2 + 128 * (j // G4) + G4 * lane + j % G4
                if p >= end:
                    break
                words.append(int(p32[p]))
                j += 1
            n = m.shape[1] // S
            syms = E.decode_stream_quad_textbook(words, n // 4, r.cdf)
            want = m[row, E.stream_columns(m.shape[1], S, lane)].astype(np.int64) - r.min_value
            assert syms == want.tolist()


def test_models():
    m = CASES["gauss6"]
    for P in (8, 12):
        mn, G, cdf = E.global_model(m, P)
        f = np.diff(cdf.astype(np.int64))
        present = np.bincount(m.astype(np.int64).ravel() - mn, minlength=G) > 0
        assert cdf[0] == 0 and cdf[-1] == 1 << P
        assert np.all(f[present] >= 1) and np.all(f[~present] == 0)
    mn, G, freq = E.per_row_model(CASES["near_const"])
    assert np.all(freq.astype(np.int64).sum(axis=1) == 256)
    row4 = freq[4]   # constant row at the bottom of the alphabet: 255 + dummy neighbour
    assert row4[0] == 255 and row4[1] == 1
    mn, G, cdf = E.global_model(CASES["const"], 12)
    assert G == 2 and cdf.tolist() == [0, 4095, 4096]
    # quantize_counts: closed forms
    assert E.quantize_counts([1, 0, 1], 256).tolist() == [128, 0, 128]
    assert E.quantize_counts([3, 1], 4).tolist() == [3, 1]


def test_ppf_and_lut_closed_form():
    cdf = np.array([0, 0, 10, 10, 200, 256])
    ppf = E.ppf_of(cdf, 8)
    want = [next(s for s in range(5) if cdf[s] <= q < cdf[s + 1]) for q in range(256)]
    assert ppf.tolist() == want
    lut = E.lut_of(cdf, -3, 8)
    q = 57
    s = want[q]
    assert lut[q] == (cdf[s + 1] - cdf[s]) | (((-3 + s) & 0xFF) << 24)
    assert lut[256 + q] == q - cdf[s]


def test_planted_column_mistake_is_caught():
    """A decoder that maps stream l to columns 16*32*(k//16) + ... for S = 64 (the
    reference's 512-column grouping) does not reproduce the matrix."""
    m = CASES["gauss6"]
    r = E.encode_record(m, 64, 8, 0)
    good = r.decode()
    bad = np.empty_like(good)
    for lane in range(64):
        k = np.arange(m.shape[1] // 64)
        col_bad = (512 * (k // 16) + 16 * lane + k % 16) % m.shape[1]
        bad[:, col_bad] = good[:, E.stream_columns(m.shape[1], 64, lane)]
    assert not np.array_equal(bad, m)


def test_container_round_trip(tmp_path):
    recs = [E.encode_record(CASES["gauss6"], 64, 12, 0), E.encode_record(CASES["near_const"], 64, 8, 1),
            E.encode_record(CASES["laplace4"], 32, 8, 0),
            E.encode_record(CASES["gauss6"], 64, 8, 0, coder=1)]
    vec = np.arange(-5, 6, dtype=np.int8)
    p = tmp_path / "a.mvx"
    E.write_mvx1(p, recs, vec, 0xDEADBEEF)
    hdr, v, back = E.read_mvx1(p)
    assert hdr == {"num_matrices": 4, "result_hash": 0xDEADBEEF} and np.array_equal(v, vec)
    for a, b in zip(recs, back):
        assert np.array_equal(a.decode(), b.decode())
    q = tmp_path / "b.mvx"
    E.write_mvx1(q, back, v, hdr["result_hash"])
    assert open(p, "rb").read() == open(q, "rb").read()


# --------------------------------------------- quad coder, streams back to back (coder 2)
@pytest.mark.parametrize("S", [32, 64, 128])
@pytest.mark.parametrize("name", sorted(CASES))
def test_lanes_round_trip(S, name):
    m = CASES[name]
    if m.shape[1] % (16 * S):
        m = m[:, : m.shape[1] // (16 * S) * 16 * S]
    if m.shape[1] == 0:
        pytest.skip("narrower than one group of the stream count")
    r = E.encode_record(m, S, 8, 0, coder=2)
    assert r.cursors[0] == E.lanes_table_u16(m.shape[0], S) and np.all(r.cursors % 8 == 0)
    assert np.array_equal(r.decode(), m)


@pytest.mark.parametrize("S", [32, 64])
def test_lanes_hold_the_quad_streams(S):
    """Coder 2 stores exactly coder 1's streams: stream l's words gathered with coder 1's
    granule formula equal those gathered from coder 2's body granules and tail, and a
    textbook decode of those words gives the stream's columns."""
    m = CASES["laplace4"][:, :2048]
    r1 = E.encode_record(m, S, 8, 0, coder=1)
    r2 = E.encode_record(m, S, 8, 0, coder=2)
    p1 = r1.payload.view(np.uint32)
    G4 = 128 // S
    assert r2.cursors[0] == E.lanes_table_u16(m.shape[0], S)
    for row in range(m.shape[0]):
        K, _, _ = E.lanes_row(r2.payload, r2.cursors, m.shape[0], S, row)
        lens = []
        for lane in range(S):
            w2 = E.lane_words(r2.payload, r2.cursors, m.shape[0], S, row, lane)
            w1 = [int(p1[int(r1.cursors[row]) // 2 + 128 * (j // G4) + G4 * lane + j % G4])
                  for j in range(len(w2))]
            assert w1 == w2
            lens.append(len(w2))
            syms = E.decode_stream_quad_textbook(w2, m.shape[1] // S // 4, r2.cdf)
            want = m[row, E.stream_columns(m.shape[1], S, lane)].astype(np.int64) - r2.min_value
            assert syms == want.tolist()
        assert K == min(lens) // 4


def test_lanes_smaller_than_granules():
    """Without the per-stream padding a row takes fewer words, most so on short rows."""
    for S, m in [(32, laplace_int8(16, 4096, 4.0, seed=5)[0]),
                 (64, laplace_int8(8, 8192, 4.0, seed=6)[0])]:
        n1 = len(E.encode_record(m, S, 8, 0, coder=1).payload)
        n2 = len(E.encode_record(m, S, 8, 0, coder=2).payload)
        assert n2 < 0.96 * n1


def test_lanes_decoder_catches_a_short_stream():
    """A planted mistake in the end offsets (one stream a word shorter) is caught: the
    decoder reads past the stream or decodes other values."""
    m = CASES["gauss6"][:, :1024]
    r = E.encode_record(m, 32, 8, 0, coder=2)
    pay = r.payload.copy()
    pay[2 * 34 + 2 + 5] -= 1   # row 2's tail end offset of stream 5
    bad = E.RecordV1(r.rows, r.cols, 32, 8, 0, r.min_value, r.G, r.cursors, pay, cdf=r.cdf,
                     coder=2)
    try:
        out = bad.decode()
    except (IndexError, ValueError):
        return
    assert not np.array_equal(out, m)
