"""Pins the family-B oracle (oracle/family_b.py) against the safetensors files written by
the reference encoder (buffered/encoder.py via tests/golden/make_fixtures.py)."""
import os

import numpy as np
import pytest

from oracle import family_b as B

from .helpers import gaussian_int8, laplace_int8

RUNS = [("256", 256, 2, 4.0), ("384_e5", 384, 1, 5.0)]


def load(golden_dir, name):
    from safetensors.numpy import load_file
    return load_file(os.path.join(golden_dir, name))


@pytest.mark.parametrize("tag,dim,count,ent", RUNS)
def test_generator_matches_raw_file(golden_dir, tag, dim, count, ent):
    raw = load(golden_dir, f"b_raw_{tag}.safetensors")
    m, v0, sc, h = B.gen_matrices(dim, count, ent)
    assert np.array_equal(m, raw["matrices"]) and np.array_equal(v0, raw["initial_vector"])
    assert np.array_equal(sc, raw["scales"]) and np.array_equal(h, raw["hashes"])


@pytest.mark.parametrize("tag,dim,count,ent", RUNS)
@pytest.mark.parametrize("layout", ["rows", "stripes"])
def test_model_and_encoder_match_file(golden_dir, tag, dim, count, ent, layout):
    raw = load(golden_dir, f"b_raw_{tag}.safetensors")
    f = load(golden_dir, f"b_{layout}_{tag}.safetensors")
    enc = B.encode_rows if layout == "rows" else B.encode_stripes
    comp, offs, size = [], [], 0
    for i, m in enumerate(raw["matrices"]):
        mn, cdf = B.fit_model(m)
        ppf, cum, prob = B.decoder_tables(mn, cdf)
        assert np.array_equal(B.pack_4bytes(ppf), f["all_ppfs"][i])
        assert np.array_equal(B.pack_4bytes(cum), f["all_ppfs_cum"][i])
        assert np.array_equal(B.pack_4bytes(prob), f["all_ppfs_prob"][i])
        c, o, _, _ = enc(m)
        comp.append(c)
        offs.append(o + size)
        size += len(c) // 4
    assert np.array_equal(np.concatenate(comp), f["compressed"])
    assert np.array_equal(np.stack(offs), f["all_warp_offsets"])


@pytest.mark.parametrize("tag,dim,count,ent", RUNS)
@pytest.mark.parametrize("layout", ["rows", "stripes"])
def test_decoder_reads_file(golden_dir, tag, dim, count, ent, layout):
    raw = load(golden_dir, f"b_raw_{tag}.safetensors")
    f = load(golden_dir, f"b_{layout}_{tag}.safetensors")
    dec = B.decode_rows if layout == "rows" else B.decode_stripes
    for i, m in enumerate(raw["matrices"]):
        mn, cdf = B.fit_model(m)
        got = dec(f["compressed"], f["all_warp_offsets"][i], B.decoder_tables(mn, cdf), dim, dim)
        assert np.array_equal(got, m)


@pytest.mark.parametrize("tag,dim,count,ent", RUNS)
def test_chain_hashes(golden_dir, tag, dim, count, ent):
    """The kernels' fp32 epilogue clamp(rint(float(acc) * scale)) reproduces the files'
    hashes (which the generator computes with float64 products)."""
    raw = load(golden_dir, f"b_raw_{tag}.safetensors")
    v = raw["initial_vector"]
    for m, s, h in zip(raw["matrices"], raw["scales"], raw["hashes"]):
        v = B.requant_b(m.astype(np.int64) @ v.astype(np.int64), s)
        assert B.fxhash(v) == h


@pytest.mark.parametrize("layout", ["rows", "stripes"])
def test_round_trips(layout):
    enc, dec = ((B.encode_rows, B.decode_rows) if layout == "rows"
                else (B.encode_stripes, B.decode_stripes))
    near = np.zeros((64, 256), np.int8)
    near[3, 7] = 2
    near[40] = -5
    for m in (gaussian_int8(64, 384, 3.0, seed=1)[0], laplace_int8(32, 512, 2.0, seed=2)[0],
              near, np.random.default_rng(3).integers(-60, 60, (32, 256)).astype(np.int8)):
        comp, offs, mn, cdf = enc(m)
        got = dec(comp, offs, B.decoder_tables(mn, cdf), m.shape[1], m.shape[0])
        assert np.array_equal(got, m)


def test_textbook_quad_pop():
    """One coder's stream decoded by four single-symbol pops per quad."""
    m = laplace_int8(1, 512, 3.0, seed=4)[0]
    mn, cdf = B.fit_model(m)
    c, buf = B.QuadCoder(), []
    quads = m[0].reshape(-1, 4)
    for q in quads[::-1]:
        c.encode(q, mn, cdf, buf)
    words = [c.state & 0xFFFFFFFF, c.state >> 32] + buf[::-1]
    assert B.quad_pop(words, len(quads), cdf) == (m[0].astype(int) - mn).tolist()


def test_model_closed_forms():
    # two symbols 3:1 -> 96 / 32 of 128; argmax adjusts the first maximum
    mn, cdf = B.fit_model(np.array([5, 5, 5, 7], np.int8))
    assert mn == 5 and cdf.tolist() == [0, 96, 96, 128]
    ppf, cum, prob = B.decoder_tables(mn, cdf)
    assert ppf[95] == 5 and ppf[96] == 7 and cum[100] == 96 and prob[100] == 32
    assert B.requant_b([3, -3, 1000], 0.5).tolist() == [2, -2, 127]   # ties to even, clamp
