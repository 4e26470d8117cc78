"""The library's source stamp: build() records a digest of the sources beside
libmatvans.so, and the loader refuses an in-tree library whose digest differs from the
current sources (edited or reverted without a rebuild)."""
import shutil

import pytest

import matvans_amd.build as B
import matvans_amd.capi as capi


def test_digest_changes_with_any_source(tmp_path, monkeypatch):
    csrc = tmp_path / "csrc"
    shutil.copytree(B.CSRC, csrc)
    monkeypatch.setattr(B, "CSRC", str(csrc))
    d0 = B.source_digest()
    assert d0 == B.source_digest()                     # deterministic
    hdr = sorted(csrc.glob("*.h"))[0]
    hdr.write_text(hdr.read_text() + "\n// edited\n")
    assert B.source_digest() != d0                     # headers count
    hip = sorted(csrc.glob("*.hip"))[0]
    hip.write_text(hip.read_text() + "\n")
    d2 = B.source_digest()
    hip.rename(hip.with_name("renamed_" + hip.name))   # names count
    assert B.source_digest() != d2


def test_built_library_matches_its_stamp():
    import os
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    capi._check_stamp()


def test_mismatched_or_missing_stamp_is_refused(tmp_path, monkeypatch):
    stamp = tmp_path / "libmatvans.sources"
    stamp.write_text("0" * 64 + "\n")
    monkeypatch.setattr(B, "STAMP", str(stamp))
    with pytest.raises(RuntimeError, match="other sources"):
        capi._check_stamp()
    stamp.unlink()
    with pytest.raises(RuntimeError, match="no source stamp"):
        capi._check_stamp()
