"""The C-ABI library loads without a GPU and exports every function include/*.h declares;
its error string is per thread."""
import ctypes
import glob
import os
import re

import matvans_amd
from matvans_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(matvans_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert "matvans_decompress_and_multiply" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    assert declared_functions() <= set(capi.SIGNATURES)


def test_version_and_error_string():
    assert matvans_amd.lib().matvans_version() >= 100
    assert isinstance(capi.last_error(), str)


def test_last_error_is_per_thread(tmp_path):
    """matvans_last_error() is thread-local: two threads that fail at the same time each
    read their own message, and a thread that has not failed reads none."""
    import threading
    L = matvans_amd.lib()
    gate = threading.Barrier(2)
    got = {}

    def fail(name):
        out = ctypes.c_void_p()
        path = str(tmp_path / f"missing_{name}.bin").encode()
        gate.wait()
        st = L.matvans_bin_read(path, ctypes.byref(out))
        gate.wait()          # both calls have failed before either message is read
        got[name] = (st, capi.last_error())

    def clean():
        got["clean"] = capi.last_error()

    ts = [threading.Thread(target=fail, args=(n,)) for n in ("a", "b")]
    ts.append(threading.Thread(target=clean))
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=60)
    for me, other in (("a", "b"), ("b", "a")):
        st, msg = got[me]
        assert st != 0
        assert f"missing_{me}.bin" in msg and f"missing_{other}.bin" not in msg
    assert got["clean"] == ""
