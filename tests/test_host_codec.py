"""Host side of the product library (no GPU): .bin reader/writer, entropy model, encoder,
FxHash and their error paths, checked against the oracle and the golden files."""
import os

import numpy as np
import pytest

import matvans_amd as mv
from matvans_amd.capi import MatvansError
from oracle import oracle as O

from .helpers import gaussian_int8, laplace_int8

FIXTURES = ["compressed_matrices_512_5bit_count2.bin", "compressed_matrices_1024_5bit_count1.bin"]


@pytest.mark.parametrize("fx", FIXTURES)
def test_reader_matches_oracle(golden_dir, fx):
    path = os.path.join(golden_dir, fx)
    hdr, vec, mats = mv.read_bin(path)
    ohdr, ovec, orecs = O.read_bin(path)
    assert hdr == ohdr and np.array_equal(vec, ovec)
    for m, r in zip(mats, orecs):
        assert (m.rows, m.cols, m.grid_spacing, m.min_value, m.G, m.payload_size) == \
               (r.rows, r.cols, r.grid_spacing, r.min_value, r.G, r.payload_size)
        assert np.array_equal(m.cursors, r.cursors)
        assert np.array_equal(m.cdf_data, r.cdf) and np.array_equal(m.ppf_data, r.ppf)
        assert np.array_equal(m.payload, r.payload)


@pytest.mark.parametrize("fx", FIXTURES)
def test_writer_reproduces_file_bytes(golden_dir, fx, tmp_path):
    path = os.path.join(golden_dir, fx)
    hdr, vec, mats = mv.read_bin(path)
    out = tmp_path / "w.bin"
    mv.write_bin(out, hdr["result_hash"], vec, mats)
    assert open(out, "rb").read() == open(path, "rb").read()


@pytest.mark.parametrize("fx", FIXTURES)
def test_model_and_encoder_match_golden(golden_dir, fx):
    hdr, vec, mats = mv.read_bin(os.path.join(golden_dir, fx))
    for m in mats:
        dense = O.c_decode(O.Record(m.rows, m.cols, 1.0, m.cursors, m.payload_size, m.min_value,
                                    m.G, m.cdf_data, m.ppf_data, m.payload))
        mn, G, cdf16, ppf = mv.build_model(dense)
        assert (mn, G) == (m.min_value, m.G)
        assert np.array_equal((cdf16.astype(np.int64) & 0xFF).astype(np.uint8), m.cdf_data)
        assert np.array_equal(ppf, m.ppf_data)
        cur, pay = mv.encode_rows(dense, mn, cdf16, n_threads=4)
        _, _, used = O.encode_matrix(dense[:4], mn, cdf16)   # padding mask, first 4 rows
        assert np.array_equal(cur, m.cursors) and len(pay) == m.payload_size
        n4 = len(used)
        assert np.array_equal(pay[:n4][used], m.payload[:n4][used])


@pytest.mark.parametrize("maker,arg,shape", [
    (gaussian_int8, 6.0, (33, 1024)), (laplace_int8, 4.0, (17, 2048)),
    (gaussian_int8, 0.3, (5, 512)), (gaussian_int8, 60.0, (9, 512))])
def test_encoder_equals_c_oracle(maker, arg, shape):
    m, _ = maker(*shape, arg, seed=11)
    mn, G, cdf16, ppf = mv.build_model(m)
    omn, oG, ocdf, _, _ = O.create_entropy_model(m)
    assert (mn, G) == (omn, oG)
    # equal counts may be ordered differently from numpy's argsort; the sum is exact
    assert cdf16[-1] == 256 and (np.diff(cdf16.astype(int)) >= 0).all()
    cur, pay = mv.encode_rows(m, mn, cdf16, n_threads=3)
    ocur, opay = O.c_encode(m, mn, cdf16)
    assert np.array_equal(cur, ocur) and np.array_equal(pay, opay)
    rec = O.Record(shape[0], shape[1], 1.0, cur, len(pay), mn, G,
                   (cdf16.astype(np.int64) & 0xFF).astype(np.uint8), ppf, pay)
    assert np.array_equal(O.c_decode(rec), m)


def test_model_equals_oracle_without_ties():
    # distinct counts everywhere: the order of equal counts cannot matter
    counts = np.arange(1, 40) ** 2
    data = np.repeat(np.arange(-19, 20), counts).astype(np.int8)
    np.random.default_rng(0).shuffle(data)
    mn, G, cdf16, ppf = mv.build_model(data)
    omn, oG, ocdf, _, _ = O.create_entropy_model(data)
    assert (mn, G) == (omn, oG) and np.array_equal(cdf16, ocdf)
    assert np.array_equal(ppf, O.create_ppf(ocdf))


def test_fxhash_matches_oracle():
    a = np.random.default_rng(5).integers(-128, 128, 10000).astype(np.int8)
    assert mv.hash_int8_array(a) == O.c_fxhash(a) == O.fxhash(a)


def test_reader_rejects_truncated_and_missing(golden_dir, tmp_path):
    data = open(os.path.join(golden_dir, FIXTURES[0]), "rb").read()
    for cut in (3, 17, 600, len(data) - 1):
        p = tmp_path / f"t{cut}.bin"
        p.write_bytes(data[:cut])
        with pytest.raises(MatvansError) as e:
            mv.read_bin(p)
        assert e.value.status == 4   # MATVANS_E_FORMAT
    with pytest.raises(MatvansError) as e:
        mv.read_bin(tmp_path / "missing.bin")
    assert e.value.status == 5       # MATVANS_E_IO


def test_reader_rejects_bad_cursor(golden_dir, tmp_path):
    data = bytearray(open(os.path.join(golden_dir, FIXTURES[0]), "rb").read())
    # first record: header 12 + len_v 4 + vec 512 -> rows/cols/gs 12 -> cursors
    off = 12 + 4 + 512 + 12 + 4 * 5
    data[off:off + 4] = (7).to_bytes(4, "little")      # cursors[5] = 7: not a granule
    p = tmp_path / "c.bin"
    p.write_bytes(bytes(data))
    with pytest.raises(MatvansError) as e:
        mv.read_bin(p)
    assert e.value.status == 4


def test_encoder_error_paths():
    m = np.zeros((2, 500), np.int8)
    with pytest.raises(MatvansError) as e:
        mv.encode_rows(m, 0, np.array([0, 255, 256], np.uint16))
    assert e.value.status == 2       # MATVANS_E_SHAPE
    m = np.full((2, 512), 9, np.int8)
    with pytest.raises(MatvansError) as e:
        mv.encode_rows(m, 0, np.array([0, 255, 256], np.uint16))
    assert e.value.status == 7       # MATVANS_E_RANGE


def test_raw_reader_matches_golden_file(golden_dir, tmp_path):
    """The library's raw-container reader (the int8 baseline's input) returns the header,
    vector and matrices the oracle reads from the golden raw file the reference generator
    wrote; every truncation is rejected with a status code."""
    import gzip
    from matvans_amd.capi import MatvansError
    data = gzip.decompress(open(os.path.join(golden_dir, "raw-matrices_512.bin.gz"), "rb").read())
    p = tmp_path / "raw-matrices_512.bin"
    p.write_bytes(data)
    hdr, vec, mats = mv.read_raw(p)
    num, rh, ovec, omats = O.read_raw_bin(os.path.join(golden_dir, "raw-matrices_512.bin.gz"))
    assert hdr == {"num_matrices": num, "result_hash": rh}
    assert np.array_equal(vec, ovec)
    assert len(mats) == len(omats)
    for a, b in zip(mats, omats):
        assert np.array_equal(a, b)
    for n in (0, 7, 11, 12 + len(vec) - 1, len(data) - 1):
        q = tmp_path / f"cut{n}.bin"
        q.write_bytes(data[:n])
        with pytest.raises(MatvansError):
            mv.read_raw(q)


def test_raw_reader_odd_matrix_and_unaligned_vector(tmp_path):
    """A raw file with a vector length that is not a multiple of 4 (padded by the writer)
    and an odd-sized matrix (one pad byte), written as create_matrices_WarpPerRow.py:94-129
    writes it."""
    import struct
    rng = np.random.default_rng(3)
    vec = rng.integers(-127, 128, 6).astype(np.int8)
    m1 = rng.integers(-127, 128, (3, 6)).astype(np.int8)
    m2 = rng.integers(-127, 128, (5, 3)).astype(np.int8)
    b = struct.pack("<LL", 2, 1234) + struct.pack("<L", 6) + vec.tobytes() + b"\0\0"
    for m in (m1, m2):
        b += struct.pack("<LLf", *m.shape, 1.0) + m.tobytes() + (b"\0" if m.size % 2 else b"")
    p = tmp_path / "odd.bin"
    p.write_bytes(b)
    hdr, v, mats = mv.read_raw(p)
    assert hdr["result_hash"] == 1234 and np.array_equal(v, vec)
    assert np.array_equal(mats[0], m1) and np.array_equal(mats[1], m2)
