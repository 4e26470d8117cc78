"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
.bin v0, MVX1 and safetensors readers on the golden files, on truncations of them
(every one must be rejected) and on seeded one-byte corruptions, plus the encoders on
edge-case matrices (tests/asan/fuzz_readers.cpp).  These readers parse files a caller
hands in, so a damaged file must end in a status code, never in an out-of-bounds access.
Built with g++ from the host sources alone (no HIP file on the line, CPU only)."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "matvans_amd" / "csrc"
GOLDEN = ROOT / "tests" / "golden"
HOST_SOURCES = ["host_codec.cpp", "host_v1.cpp", "host_b.cpp"]


@pytest.fixture(scope="module")
def fuzz_binary(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("asan") / "fuzz_readers"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "asan" / "fuzz_readers.cpp"),
           *[str(CSRC / s) for s in HOST_SOURCES], "-o", str(out), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode and "asan" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("sanitizer runtime not installed")
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def test_readers_and_encoders_under_sanitizers(fuzz_binary, tmp_path):
    files = sorted(GOLDEN.glob("*.bin")) + sorted(GOLDEN.glob("*.safetensors"))
    assert files, "golden files missing"
    # the raw baseline file, unpacked (the reader takes plain files)
    import gzip
    raw = tmp_path / "raw-matrices_512.raw"
    raw.write_bytes(gzip.decompress((GOLDEN / "raw-matrices_512.bin.gz").read_bytes()))
    files.append(raw)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(fuzz_binary), str(tmp_path), *map(str, files)],
                       capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "fuzz ok" in r.stdout
