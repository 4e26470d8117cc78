"""The C++ host driver (matvans_amd/lib/matvans_run): the reference's chained benchmark
harness (matvansWarpPerRow.cu:225-393) over the C ABI.  CPU tests check the binary and its
argument / error handling; GPU tests run the chain on the golden .bin files and on MVX1
files re-encoded from the golden raw matrices, and check the FxHash of the last vector
against the header's result_hash."""
import json
import os
import subprocess

import pytest

from matvans_amd.build import DRIVER

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def run(*args, timeout=120):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)


def test_driver_built_and_usage():
    assert os.access(DRIVER, os.X_OK)
    r = run("--help")
    assert r.returncode == 0 and "usage: matvans_run" in r.stdout


def test_driver_rejects_missing_file(tmp_path):
    r = run(str(tmp_path / "nope.bin"))
    assert r.returncode == 2 and "read .bin" in r.stderr


def test_driver_rejects_truncated_file(tmp_path):
    src = open(os.path.join(GOLDEN, "compressed_matrices_512_5bit_count2.bin"), "rb").read()
    p = tmp_path / "cut.bin"
    p.write_bytes(src[: len(src) // 2])
    r = run(str(p))
    assert r.returncode == 2 and "read .bin" in r.stderr


def last_json(out):
    return json.loads(out.strip().splitlines()[-1])


@pytest.fixture
def gpu_ready(cuda):
    """The test process initialises its GPU runtime (one small device call through the
    library, which shares torch's HIP runtime) before any child process uses the GPU.
    The failures once seen when children ran first came from this process mapping two
    HIP runtimes (library loaded before torch); capi.lib() now maps torch's first."""
    import torch
    import matvans_amd as mv
    w = torch.ones((4, 16), dtype=torch.int8, device=cuda)
    x = torch.ones((1, 16), dtype=torch.int8, device=cuda)
    assert int(mv.int8_matvec(w, x)[0, 0]) == 16
    torch.cuda.synchronize()


@pytest.mark.gpu
@pytest.mark.parametrize("fx", ["compressed_matrices_512_5bit_count2.bin",
                                "compressed_matrices_1024_5bit_count1.bin"])
def test_driver_chain_on_golden_bin(gpu_ready, fx):
    r = run(os.path.join(GOLDEN, fx), "--iters", "3")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Hashes match!") == 3
    j = last_json(r.stdout)
    assert j["hash_matches"] == 3 and j["container"] == "bin v0"


@pytest.mark.gpu
@pytest.mark.parametrize("streams,precision,coder,outliers",
                         [(32, 8, 1, None), (32, 12, 0, None), (32, 8, 0, None), (32, 8, 1, 3)])
def test_driver_chain_on_mvx1(gpu_ready, tmp_path, streams, precision, coder, outliers):
    """The golden raw matrices re-encoded as MVX1: same chain, same result_hash; with a
    narrow outlier band the file is a version-2 container whose outlier lists the driver
    reads, uploads and applies."""
    from matvans_amd.mvx import CompressedMatrixV1, write_mvx1
    from oracle import oracle as O
    num, result_hash, vec, mats = O.read_raw_bin(os.path.join(GOLDEN, "raw-matrices_512.bin.gz"))
    enc = [CompressedMatrixV1.encode(m, streams, precision, 0, coder=coder, outliers=outliers)
           for m in mats]
    if outliers is not None:
        assert all(e.n_outliers > 0 for e in enc)
    p = tmp_path / "chain.mvx"
    write_mvx1(p, enc, vector=vec, result_hash=result_hash)
    r = run(str(p), "--iters", "2", "--quiet")
    assert r.returncode == 0, r.stdout + r.stderr
    j = last_json(r.stdout)
    assert j["hash_matches"] == 2 and j["container"] == "MVX1"


def test_driver_raw_usage_and_missing_file(tmp_path):
    r = run("--raw", str(tmp_path / "nope.bin"))
    assert r.returncode == 2 and "read raw file" in r.stderr
    r = run("--help")
    assert "--raw" in r.stdout


@pytest.mark.gpu
def test_driver_raw_int8_chain_reproduces_hash(gpu_ready, tmp_path):
    """The int8 comparator's harness (baseline_matVec_custom.cu:434-620) on the golden raw
    file: int8 matvec + clamped requant per matrix, FxHash of the last vector equals the
    header's result_hash."""
    import gzip
    p = tmp_path / "raw-matrices_512.bin"
    p.write_bytes(gzip.decompress(open(os.path.join(GOLDEN, "raw-matrices_512.bin.gz"),
                                       "rb").read()))
    r = run("--raw", str(p), "--iters", "3")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Hashes match!") == 3
    j = last_json(r.stdout)
    assert j["hash_matches"] == 3 and j["container"] == "raw"


@pytest.mark.gpu
def test_python_raw_int8_chain_reproduces_hash(gpu_ready, tmp_path):
    """The same chain through the Python bindings: library raw reader, int8 comparator,
    clamped device requant."""
    import gzip
    import torch
    import matvans_amd as mv
    p = tmp_path / "raw-matrices_512.bin"
    p.write_bytes(gzip.decompress(open(os.path.join(GOLDEN, "raw-matrices_512.bin.gz"),
                                       "rb").read()))
    hdr, vec, mats = mv.read_raw(p)
    v = torch.from_numpy(vec).cuda()
    for m in mats:
        acc = mv.int8_matvec(torch.from_numpy(m).cuda(), v)
        v, _ = mv.absmax_requant(acc, clamp=True)
    assert mv.hash_int8_array(v.cpu().numpy()) == hdr["result_hash"]
