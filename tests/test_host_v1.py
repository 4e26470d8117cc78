"""Host side of the MVX1 extensions in the product library (no GPU): the C++ model,
encoder, decode table and container against the oracle, and their error paths."""
import numpy as np
import pytest

from matvans_amd.capi import MatvansError
from matvans_amd.mvx import CompressedMatrixV1, read_mvx1, write_mvx1
from oracle import ext as E

from .helpers import gaussian_int8, laplace_int8

VARIANTS = [(32, 8, 0), (64, 8, 0), (128, 8, 0), (32, 12, 0), (64, 12, 0), (128, 12, 0),
            (64, 8, 1)]


def _mats():
    near = np.zeros((5, 2048), np.int8)
    near[2, 100] = -4
    near[3] = 127
    return [gaussian_int8(6, 2048, 6.0, seed=11)[0], laplace_int8(5, 4096, 4.0, seed=12)[0],
            near, np.random.default_rng(5).integers(-128, 128, (3, 2048)).astype(np.int8)]


@pytest.mark.parametrize("S,P,mode", VARIANTS)
def test_encoder_matches_oracle(S, P, mode):
    for m in _mats():
        c = CompressedMatrixV1.encode(m, S, P, mode, n_threads=3)
        r = E.encode_record(m, S, P, mode)
        assert (c.min_value, c.G) == (r.min_value, r.G)
        assert np.array_equal(c.cursors, r.cursors)
        assert np.array_equal(c.payload, r.payload)
        if mode == 0:
            assert np.array_equal(c.cdf, r.cdf)
            assert np.array_equal(c.lut(), E.lut_of(r.cdf, r.min_value, P))
        else:
            assert np.array_equal(c.freq, r.freq)
        assert np.array_equal(r.decode(), m)


@pytest.mark.parametrize("S,mode", [(32, 0), (64, 0), (128, 0), (64, 1)])
def test_quad_encoder_matches_oracle(S, mode):
    for m in _mats():
        if m.shape[1] % (16 * S):
            continue
        c = CompressedMatrixV1.encode(m, S, 8, mode, n_threads=3, coder=1)
        r = E.encode_record(m, S, 8, mode, coder=1)
        assert np.array_equal(c.cursors, r.cursors) and np.array_equal(c.payload, r.payload)
        if mode == 0:
            assert np.array_equal(c.cdf, r.cdf)
        else:
            assert np.array_equal(c.freq, r.freq)
        assert np.array_equal(r.decode(), m)


def test_reference_special_case_matches_v0_library():
    import matvans_amd as mv
    m = laplace_int8(9, 3072, 4.0, seed=4)[0]
    v0 = mv.CompressedMatrix.encode(m)
    v1 = CompressedMatrixV1.encode(m, 32, 8, 0)
    assert np.array_equal(v0.cursors, v1.cursors) and np.array_equal(v0.payload, v1.payload)


def test_container_round_trip(tmp_path):
    ms = _mats()
    cs = [CompressedMatrixV1.encode(ms[0], 64, 12, 0), CompressedMatrixV1.encode(ms[2], 64, 8, 1),
          CompressedMatrixV1.encode(ms[1], 32, 8, 0), CompressedMatrixV1.encode(ms[1], 64, coder=1),
          CompressedMatrixV1.encode(ms[2], 64, 8, 1, coder=1)]
    vec = np.arange(7, dtype=np.int8)
    p = tmp_path / "a.mvx"
    write_mvx1(p, cs, vec, 77)
    hdr, v, back = read_mvx1(p)
    assert hdr == {"num_matrices": 5, "result_hash": 77} and np.array_equal(v, vec)
    ohdr, ov, orecs = E.read_mvx1(p)
    for c, b, o in zip(cs, back, orecs):
        assert np.array_equal(c.payload, b.payload) and np.array_equal(c.cursors, b.cursors)
        assert np.array_equal(o.decode(), b_mat(c)) and b.coder == c.coder
    q = tmp_path / "b.mvx"
    E.write_mvx1(q, orecs, ov, ohdr["result_hash"])
    assert open(p, "rb").read() == open(q, "rb").read()


def b_mat(c):
    return E.RecordV1(c.rows, c.cols, c.streams, c.precision, c.table_mode, c.min_value, c.G,
                      c.cursors, c.payload, c.cdf, c.freq, coder=c.coder).decode()


def test_reader_rejects_bad_files(tmp_path):
    c = CompressedMatrixV1.encode(_mats()[0], 64, 8, 0)
    p = tmp_path / "a.mvx"
    write_mvx1(p, [c])
    data = open(p, "rb").read()
    for name, blob in [("magic", b"MVX2" + data[4:]), ("truncated", data[:-6]),
                       ("trailing", data + b"\0\0\0\0"),
                       ("streams", data[:32] + (48).to_bytes(4, "little") + data[36:])]:
        q = tmp_path / f"{name}.mvx"
        open(q, "wb").write(blob)
        with pytest.raises(MatvansError) as e:
            read_mvx1(q)
        assert e.value.status == 4, name


def test_encoder_argument_errors():
    m = _mats()[0]
    with pytest.raises(MatvansError) as e:       # 12-bit per-row tables
        CompressedMatrixV1.encode(m, 64, 12, 1)
    assert e.value.status == 8
    with pytest.raises(MatvansError) as e:       # cols not a multiple of 16 * 64
        CompressedMatrixV1.encode(m[:, :512], 64, 8, 0)
    assert e.value.status == 2
    with pytest.raises(MatvansError) as e:       # quad coder with 12-bit tables
        CompressedMatrixV1.encode(m, 64, 12, 0, coder=1)
    assert e.value.status == 8


# ------------------------------------------------------------------ outlier split
def test_outlier_split_matches_oracle_and_round_trips():
    """The library's split equals the oracle's, and the quad encoding of the coded matrix
    plus the records gives back W (oracle decode + scatter)."""
    for m in [laplace_int8(7, 4096, 4.0, seed=21)[0], gaussian_int8(5, 2048, 6.0, seed=22)[0]]:
        band = CompressedMatrixV1.outlier_band(m)
        assert 0 < band < 127
        coded, start, rec = CompressedMatrixV1.split_outliers(m, band)
        ocoded, ostart, orec = E.split_outliers(m, band)
        assert np.array_equal(coded, ocoded) and np.array_equal(start, ostart)
        assert np.array_equal(rec, orec)
        assert np.abs(coded.astype(np.int16)).max() <= band
        c = CompressedMatrixV1.encode(m, 32, 8, 0, coder=1, outliers=band)
        assert c.n_outliers == rec.size > 0
        r = E.encode_record(coded, 32, 8, 0, coder=1)
        assert np.array_equal(c.payload, r.payload) and np.array_equal(c.cdf, r.cdf)
        assert np.array_equal(E.merge_outliers(r.decode(), c.out_start, c.out_rec), m)


def test_outlier_split_pays_on_a_long_tail_and_not_without_one():
    """Laplace(0, 4): the split stores fewer bytes than the plain quad encoding; a
    distribution without a tail (uniform over 17 values) gets band 127 (no split)."""
    m = laplace_int8(64, 4096, 4.0, seed=23)[0]
    plain = CompressedMatrixV1.encode(m, 32, 8, 0, coder=1)
    split = CompressedMatrixV1.encode(m, 32, 8, 0, coder=1, outliers="auto")
    assert split.n_outliers > 0
    assert split.stored_bytes < 0.93 * plain.stored_bytes
    u = np.random.default_rng(3).integers(-8, 9, (16, 2048)).astype(np.int8)
    assert CompressedMatrixV1.outlier_band(u) == 127
    assert CompressedMatrixV1.encode(u, 32, 8, 0, coder=1, outliers="auto").n_outliers == 0


def test_outlier_split_extremes():
    """band 0: every nonzero entry is an outlier (the streams code an all-zero matrix);
    a split of a matrix without entries beyond the band lists nothing."""
    m = gaussian_int8(3, 1024, 6.0, seed=24)[0]
    coded, start, rec = CompressedMatrixV1.split_outliers(m, 0)
    assert not coded.any() and rec.size == np.count_nonzero(m) and start[-1] == rec.size
    assert np.array_equal(E.merge_outliers(coded, start, rec), m)
    coded, start, rec = CompressedMatrixV1.split_outliers(m, 127)
    assert np.array_equal(coded, m) and rec.size == 0 and not start.any()
    with pytest.raises(MatvansError):
        CompressedMatrixV1.split_outliers(m, 200)


def test_outlier_container_round_trip(tmp_path):
    """MVX1 version 2: the library writes the outlier records, reads them back, and the
    oracle's reader sees the same records; a file without outliers stays version 1."""
    import struct
    m = laplace_int8(9, 2048, 4.0, seed=25)[0]
    c = CompressedMatrixV1.encode(m, 32, 8, 0, coder=1, outliers="auto")
    plain = CompressedMatrixV1.encode(m, 64, 8, 0, coder=1)
    p = tmp_path / "o.mvx"
    write_mvx1(p, [c, plain], vector=np.arange(5, dtype=np.int8), result_hash=7)
    assert struct.unpack_from("<L", p.read_bytes(), 4)[0] == 2
    hdr, vec, mats = read_mvx1(p)
    assert mats[0].n_outliers == c.n_outliers and mats[1].n_outliers == 0
    assert np.array_equal(mats[0].out_start, c.out_start)
    assert np.array_equal(mats[0].out_rec, c.out_rec)
    _, _, recs = E.read_mvx1(p)
    assert np.array_equal(E.merge_outliers(recs[0].decode(), recs[0].out_start,
                                           recs[0].out_rec), m)
    assert np.array_equal(recs[1].decode(), m)
    q = tmp_path / "plain.mvx"
    write_mvx1(q, [plain])
    assert struct.unpack_from("<L", q.read_bytes(), 4)[0] == 1
    # truncated outlier section: refused
    data = p.read_bytes()
    r = tmp_path / "cut.mvx"
    r.write_bytes(data[:-(4 * plain.stored_bytes // 8)])
    with pytest.raises(MatvansError):
        read_mvx1(r)


# --------------------------------------------- quad coder, streams back to back (coder 2)
@pytest.mark.parametrize("S", [32, 64, 128])
def test_lanes_encoder_matches_oracle(S):
    for m in _mats():
        c = CompressedMatrixV1.encode(m, S, 8, 0, n_threads=3, coder=2)
        r = E.encode_record(m, S, 8, 0, coder=2)
        assert np.array_equal(c.cursors, r.cursors) and np.array_equal(c.payload, r.payload)
        assert np.array_equal(c.cdf, r.cdf)
        assert np.array_equal(r.decode(), m)
        # the Gaussian and Laplace matrices (near-constant and uniform rows leave coder 1
        # little padding, and the offsets cost more there)
        if S <= 64 and 4 < len(np.unique(m)) < 200:
            c1 = CompressedMatrixV1.encode(m, S, 8, 0, coder=1)
            assert c.payload_size < c1.payload_size


def test_lanes_container_round_trip_with_outliers(tmp_path):
    m = laplace_int8(6, 4096, 4.0, seed=21)[0]
    cs = [CompressedMatrixV1.encode(m, 32, coder=2, outliers="auto"),
          CompressedMatrixV1.encode(m, 64, coder=2)]
    assert cs[0].n_outliers > 0
    p = tmp_path / "l.mvx"
    write_mvx1(p, cs)
    _, _, back = read_mvx1(p)
    _, _, orecs = E.read_mvx1(p)
    for c, b, o in zip(cs, back, orecs):
        assert b.coder == 2 and np.array_equal(c.payload, b.payload)
        assert np.array_equal(c.cursors, b.cursors) and b.n_outliers == c.n_outliers
        w = o.decode()
        if c.n_outliers:
            w = E.merge_outliers(w, c.out_start, c.out_rec)
        assert np.array_equal(w, m)


def test_lanes_reader_rejects_bad_offsets(tmp_path):
    c = CompressedMatrixV1.encode(_mats()[0], 32, 8, 0, coder=2)
    t = 34   # row 1's table: K (two u16), then 32 tail end offsets
    for name, edit in [("decreasing", lambda pay: pay.__setitem__(t + 2 + 3, 0)),
                       ("overrun", lambda pay: pay.__setitem__(t + 2 + 31, 0xFFFF)),
                       ("granules", lambda pay: pay.__setitem__(t, pay[t] + 50))]:
        pay = c.payload.copy()
        edit(pay)
        bad = CompressedMatrixV1(c.rows, c.cols, 32, 8, 0, c.min_value, c.G, c.cursors, pay,
                                 cdf=c.cdf, coder=2)
        p = tmp_path / f"{name}.mvx"
        write_mvx1(p, [bad])
        with pytest.raises(MatvansError) as e:
            read_mvx1(p)
        assert e.value.status == 4, name


def test_lanes_tails_beyond_16bit_offsets_are_refused():
    """The tails' end offsets are u16: a row whose stream 0 is much shorter than the others
    (its columns constant, the rest uniform) leaves > 65535 tail words and is refused."""
    n = 1 << 21
    m = np.random.default_rng(3).integers(-64, 64, (1, n)).astype(np.int8)
    m[0, (np.arange(n) % 512) < 16] = 0            # stream 0's columns
    with pytest.raises(MatvansError) as e:
        CompressedMatrixV1.encode(m, 32, 8, 0, coder=2)
    assert e.value.status == 2
    ok = CompressedMatrixV1.encode(m[:, : 1 << 16], 32, 8, 0, coder=2)   # short rows fit
    assert np.array_equal(E.RecordV1(1, 1 << 16, 32, 8, 0, ok.min_value, ok.G, ok.cursors,
                                     ok.payload, cdf=ok.cdf, coder=2).decode(), m[:, : 1 << 16])


def test_lanes_need_one_table_and_8_bits():
    m = _mats()[0]
    for kw in [dict(precision=12), dict(table_mode=1)]:
        with pytest.raises(MatvansError) as e:
            CompressedMatrixV1.encode(m, 64, coder=2, **{"precision": 8, "table_mode": 0, **kw})
        assert e.value.status == 8
