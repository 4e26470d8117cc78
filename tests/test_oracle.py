"""Pins the CPU oracle (oracle/) against the reference's golden outputs and closed forms.

The golden files were written by the reference generator
(matvansWarpPerRow/create_matrices_WarpPerRow.py, see tests/golden/make_fixtures.py).
"""
import io
import os

import numpy as np
import pytest

from oracle import oracle as O

from .helpers import gaussian_int8, laplace_int8

FX512 = "compressed_matrices_512_5bit_count2.bin"
FX1024 = "compressed_matrices_1024_5bit_count1.bin"
RAW512 = "raw-matrices_512.bin.gz"


@pytest.fixture(scope="module")
def fx512(golden_dir):
    return O.read_bin(os.path.join(golden_dir, FX512))


@pytest.fixture(scope="module")
def fx1024(golden_dir):
    return O.read_bin(os.path.join(golden_dir, FX1024))


@pytest.fixture(scope="module")
def raw512(golden_dir):
    return O.read_raw_bin(os.path.join(golden_dir, RAW512))


@pytest.fixture(scope="module", autouse=True)
def c_oracle():
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle")
    if not os.path.exists(os.path.join(here, "liboracle.so")):
        subprocess.run(["make", "-s", "-C", here], check=True)


# ------------------------------------------------------------------------- hash
def test_fxhash_known_answers():
    # h = rotl(h,5) ^ b, h *= 0x27220A95: closed forms for short inputs
    assert O.fxhash(np.zeros(0, np.int8)) == 0
    assert O.fxhash(np.array([1], np.int8)) == 0x27220A95
    h1 = 0x27220A95
    h2 = ((((h1 << 5) | (h1 >> 27)) & 0xFFFFFFFF) ^ 0xFF) * 0x27220A95 & 0xFFFFFFFF
    assert O.fxhash(np.array([1, -1], np.int8)) == h2
    a = np.random.default_rng(0).integers(-128, 128, 5000).astype(np.int8)
    assert O.c_fxhash(a) == O.fxhash(a)


# ------------------------------------------------------------------- generator
def test_generator_matches_golden_raw(raw512):
    num, rh, vec, mats = raw512
    q, wd, v = O.gen_reference_matrices(512, 2, 5)
    assert num == 2
    assert np.array_equal(v, vec)
    for i in range(2):
        assert np.array_equal(q[i], mats[i])


def test_chain_hash_matches_headers(fx512, fx1024, raw512):
    _, rh, _, _ = raw512
    q, wd, v = O.gen_reference_matrices(512, 2, 5)
    assert O.reference_chain(q, wd, v)[1] == rh == fx512[0]["result_hash"]
    q, wd, v = O.gen_reference_matrices(1024, 1, 5)
    assert O.reference_chain(q, wd, v)[1] == fx1024[0]["result_hash"]


def test_fp32_device_chain_matches_header(fx512, raw512, fx1024):
    """The reference's device chain (int32 acc, fp32 absmax/127, rint) reproduces the
    float64-chain header hash on both golden files."""
    _, rh, vec, mats = raw512
    v = vec
    for m in mats:
        v, _ = O.requant_f32(O.matvec_i32(m, v))
    assert O.fxhash(v) == rh
    q, wd, v0 = O.gen_reference_matrices(1024, 1, 5)
    v, _ = O.requant_f32(O.matvec_i32(q[0], v0))
    assert O.fxhash(v) == fx1024[0]["result_hash"]


def test_requant_closed_form():
    acc = np.array([0, 1, -1, 254, -254, 127, 2**24 + 1, -(2**24)], np.int32)
    y, d = O.requant_f32(acc)
    a = np.float32(2**24 + 1)          # float(max|acc|) rounds to 2^24
    assert d == np.float32(a / np.float32(127))
    assert y[6] == 127 and y[7] == -127
    # ties round to even: acc = 127 * k + 63.5 * delta cannot occur for integers when
    # delta = 2; check a hand-made half case
    y, d = O.requant_f32(np.array([254, 1, 3, 5], np.int32))   # delta = 2
    assert d == np.float32(2.0)
    assert list(y) == [127, 0, 2, 2]   # 0.5 -> 0, 1.5 -> 2, 2.5 -> 2


def test_normalize_round_f32_half_away_and_closed_form():
    """The fp32 chain's epilogue: delta = max|y| / 127, roundf(y / delta) with ties away
    from zero (np.rint would give 2 for 2.5, 0 for 0.5), against floor(|q| + 0.5) taken
    in float64 (exact for float32 q) on random data."""
    y = np.array([127.0, 0.5, -0.5, 1.5, 2.5, -2.5, 0.49999997, -126.5, 3.0], np.float32)
    r, d = O.normalize_round_f32(y)
    assert d == np.float32(1.0)
    assert list(r) == [127, 1, -1, 2, 3, -3, 0, -127, 3]
    rng = np.random.default_rng(4)
    for n in (1, 1000, 14336):
        y = (rng.standard_normal(n) * rng.choice([1e-3, 1.0, 3e4])).astype(np.float32)
        r, d = O.normalize_round_f32(y)
        assert d == np.float32(np.abs(y).max()) / np.float32(127)
        q = (y / d).astype(np.float32).astype(np.float64)
        want = np.sign(q) * np.floor(np.abs(q) + 0.5)
        assert np.array_equal(r.astype(np.float64), want)


# --------------------------------------------------------------- model and ppf
@pytest.mark.parametrize("which", [0, 1])
def test_model_and_ppf_match_golden(fx512, raw512, which):
    _, _, _, mats = raw512
    rec = fx512[2][which]
    mn, G, cdf, _e, _x = O.create_entropy_model(mats[which])
    assert mn == rec.min_value and G == rec.G
    assert np.array_equal((cdf.astype(np.int64) & 0xFF).astype(np.uint8), rec.cdf)
    assert np.array_equal(O.create_ppf(cdf), rec.ppf)


def test_ppf_closed_form():
    cdf = np.array([0, 10, 10, 200, 256], np.uint16)   # symbol 1 has probability 0
    ppf = O.create_ppf(cdf)
    assert list(np.unique(ppf)) == [0, 2, 3]
    assert (ppf[:10] == 0).all() and (ppf[10:200] == 2).all() and (ppf[200:] == 3).all()


def test_model_special_cases():
    mn, G, cdf, e, _ = O.create_entropy_model(np.full((4, 4), -3, np.int8))
    assert (mn, G, list(cdf), e) == (-3, 2, [0, 255, 256], 0.0)
    d = np.array([0] * 1000 + [5], np.int8)    # rare value gets the minimum 1/256
    mn, G, cdf, _, _ = O.create_entropy_model(d)
    assert mn == 0 and G == 6 and cdf[1] == 255 and cdf[-1] == 256
    assert (np.diff(cdf.astype(int)) >= 0).all()


# --------------------------------------------------------------------- decode
@pytest.mark.parametrize("which", [0, 1])
def test_decode_golden_512(fx512, raw512, which):
    rec = fx512[2][which]
    mats = raw512[3]
    assert np.array_equal(rec.decode(), mats[which])
    assert np.array_equal(O.c_decode(rec), mats[which])


def test_decode_golden_1024(fx1024):
    rec = fx1024[2][0]
    q, _, _ = O.gen_reference_matrices(1024, 1, 5)
    # two granules per row and two 512-column groups per lane are exercised
    assert max(np.diff(rec.cursors.astype(np.int64))) == 512
    assert np.array_equal(rec.decode(), q[0])
    assert np.array_equal(O.c_decode(rec), q[0])


def test_planted_column_mapping_mistake_is_caught(fx1024, monkeypatch):
    """A transposed column mapping (lane-major instead of 512-column groups) decodes the
    golden file to a different matrix."""
    rec = fx1024[2][0]
    good = rec.decode()
    orig = O.GROUP_COLS
    monkeypatch.setattr(O, "GROUP_COLS", 256)
    try:
        bad = O.decode_matrix(rec.rows, rec.cols // 2 * 2, rec.cursors, rec.payload,
                              rec.min_value, rec.cdf, rec.ppf)
    except (IndexError, ValueError):
        bad = None
    monkeypatch.setattr(O, "GROUP_COLS", orig)
    assert bad is None or not np.array_equal(bad, good)


# --------------------------------------------------------------------- encode
@pytest.mark.parametrize("fx,which", [(FX512, 0), (FX512, 1), (FX1024, 0)])
def test_encode_matches_golden_outside_padding(golden_dir, fx, which):
    hdr, vec, recs = O.read_bin(os.path.join(golden_dir, fx))
    rec = recs[which]
    m = rec.decode()
    mn, G, cdf, _, _ = O.create_entropy_model(m)
    cur, pay, used = O.encode_matrix(m, mn, cdf)
    assert np.array_equal(cur, rec.cursors)
    assert len(pay) == rec.payload_size
    assert np.array_equal(pay[used], rec.payload[used])
    ccur, cpay = O.c_encode(m, mn, cdf)
    assert np.array_equal(ccur, cur) and np.array_equal(cpay, pay)


def test_bin_byte_round_trip(golden_dir, tmp_path):
    for fx in (FX512, FX1024):
        src = os.path.join(golden_dir, fx)
        hdr, vec, recs = O.read_bin(src)
        out = tmp_path / fx
        O.write_bin(out, hdr["result_hash"], vec, recs, hdr["max_word_count"])
        assert open(src, "rb").read() == open(out, "rb").read()


# ------------------------------------------------------------ closed forms
@pytest.mark.parametrize("maker,arg,shape", [
    (gaussian_int8, 6.0, (5, 1024)),
    (gaussian_int8, 0.4, (3, 512)),     # near-constant rows (few distinct values)
    (gaussian_int8, 40.0, (4, 1536)),   # wide alphabet, clipped tails
    (laplace_int8, 4.0, (6, 2048)),
])
def test_round_trip_numpy_and_c(maker, arg, shape):
    m, _ = maker(*shape, arg, seed=7)
    mn, G, cdf, _, _ = O.create_entropy_model(m)
    cur, pay = O.c_encode(m, mn, cdf)
    rec = O.Record(shape[0], shape[1], 1.0, cur, len(pay), mn, G,
                   (cdf.astype(np.int64) & 0xFF).astype(np.uint8), O.create_ppf(cdf), pay)
    assert np.array_equal(rec.decode(), m)
    assert np.array_equal(O.c_decode(rec), m)
    v = np.random.default_rng(1).integers(-127, 128, shape[1]).astype(np.int8)
    assert np.array_equal(O.c_matvec(rec, v, 3), O.matvec_i32(m, v))


def test_constant_and_extreme_matrices():
    for val in (-127, 0, 127):
        m = np.full((2, 512), val, np.int8)
        mn, G, cdf, _, _ = O.create_entropy_model(m)
        cur, pay = O.c_encode(m, mn, cdf)
        rec = O.Record(2, 512, 1.0, cur, len(pay), mn, G,
                       (cdf.astype(np.int64) & 0xFF).astype(np.uint8), O.create_ppf(cdf), pay)
        assert np.array_equal(O.c_decode(rec), m)


def test_python_coder_round_trip_small():
    """AnsCoder.push against the textbook inverse on a tiny alphabet."""
    cdf = np.array([0, 1, 128, 255, 256], np.uint16)
    rng = np.random.default_rng(3)
    syms = rng.integers(0, 4, 200)
    c = O.AnsCoder()
    for s in syms[::-1]:
        c.push(int(s), cdf)
    c.interrupt()
    stream = c.bulk[::-1]
    x = (stream[0] << 16) | stream[1]
    pos = 2
    ppf = O.create_ppf(cdf)
    out = []
    for _ in syms:
        q = x & 0xFF
        s = int(ppf[q])
        out.append(s)
        x = (x >> 8) * (int(cdf[s + 1]) - int(cdf[s])) + q - int(cdf[s])
        if x < 1 << 16:
            x = (x << 16) | stream[pos]
            pos += 1
    assert out == list(syms) and pos == len(stream) and x == 1 << 16


def test_clamped_requant_chain_reproduces_raw_file_hash(golden_dir):
    """The int8 baseline's chain (int32 matvec, absmax/127, clamped rint; baseline_matVec_
    custom.cu:64-86,399-430) over the golden raw file reproduces its header result_hash
    (:574-586), and the clamp never changes a value: |float(acc) / (max/127)| <= 127."""
    num, rh, vec, mats = O.read_raw_bin(os.path.join(golden_dir, "raw-matrices_512.bin.gz"))
    v = vec
    for m in mats:
        acc = O.matvec_i32(m, v)
        v, _ = O.requant_f32(acc, clamp=True)
        assert np.array_equal(v, O.requant_f32(acc)[0])
    assert O.fxhash(v) == rh
    rng = np.random.default_rng(9)
    for acc in (rng.integers(-2**31 + 1, 2**31 - 1, 5000).astype(np.int32),
                np.array([2**31 - 1, -2**31 + 1, 7], np.int32)):
        assert np.array_equal(O.requant_f32(acc, clamp=True)[0], O.requant_f32(acc)[0])
