"""Host side of the family-B path in the product library (no GPU): the safetensors
reader, the 7-bit model and the rows / stripes encoders against the reference's files
and the oracle, and their error paths."""
import os

import numpy as np
import pytest

from matvans_amd.buffered import BufferedMatrices, read_safetensors
from matvans_amd.capi import MatvansError
from oracle import family_b as B

from .helpers import gaussian_int8, laplace_int8

FILES = ["b_rows_256", "b_stripes_256", "b_raw_256", "b_rows_384_e5", "b_stripes_384_e5",
         "b_raw_384_e5"]


@pytest.mark.parametrize("name", FILES)
def test_reader_matches_safetensors(golden_dir, name):
    from safetensors.numpy import load_file
    path = os.path.join(golden_dir, name + ".safetensors")
    a, b = read_safetensors(path), load_file(path)
    assert sorted(a) == sorted(b)
    for k in b:
        assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("tag", ["256", "384_e5"])
@pytest.mark.parametrize("layout", ["rows", "stripes"])
def test_encoder_reproduces_reference_files(golden_dir, tag, layout):
    raw = read_safetensors(os.path.join(golden_dir, f"b_raw_{tag}.safetensors"))
    f = BufferedMatrices.read(os.path.join(golden_dir, f"b_{layout}_{tag}.safetensors"))
    assert f.layout == layout
    bm = BufferedMatrices.encode(list(raw["matrices"]), layout)
    assert np.array_equal(bm.compressed, f.compressed)
    assert all(np.array_equal(a, b) for a, b in zip(bm.offsets, f.offsets))
    assert np.array_equal(bm.ppfs, f.ppfs) and np.array_equal(bm.cums, f.cums)
    assert np.array_equal(bm.probs, f.probs)


@pytest.mark.parametrize("layout", ["rows", "stripes"])
def test_encoder_matches_oracle_on_other_shapes(layout):
    enc = B.encode_rows if layout == "rows" else B.encode_stripes
    for m in (gaussian_int8(64, 384, 3.0, seed=5)[0], laplace_int8(32, 512, 6.0, seed=6)[0],
              np.full((32, 128), -7, np.int8)):
        bm = BufferedMatrices.encode([m], layout)
        c, o, mn, cdf = enc(m)
        assert np.array_equal(bm.compressed, c) and np.array_equal(bm.offsets[0], o)


def test_model_rejects_wide_alphabet():
    """More than 128 distinct values: the reference divides by zero (encoder.py:221)."""
    m = np.tile(np.arange(-100, 100, dtype=np.int8), (32, 1))[:, :196]
    with pytest.raises(MatvansError) as e:
        BufferedMatrices.encode([m], "stripes")
    assert e.value.status == 7


def test_reader_rejects_bad_files(tmp_path, golden_dir):
    data = open(os.path.join(golden_dir, "b_rows_256.safetensors"), "rb").read()
    for name, blob in [("short", data[:5]), ("header", (10 ** 9).to_bytes(8, "little") + data[8:]),
                       ("json", data[:8] + b"[" + data[9:])]:
        p = tmp_path / f"{name}.safetensors"
        open(p, "wb").write(blob)
        with pytest.raises(MatvansError) as e:
            read_safetensors(p)
        assert e.value.status == 4, name
