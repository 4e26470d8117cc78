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


def test_backward_refuses_partial_sums_beyond_4gib():
    """The backward's group store addresses the partial sums with 32-bit byte offsets; a
    matrix whose eight planes of 64-row partial sums would pass 4 GiB is refused with a
    shape error by the argument checks, before anything reaches the GPU (the pointers
    below are never dereferenced)."""
    L = matvans_amd.lib()
    fake = ctypes.c_void_p(256)
    rows, cols = 1 << 24, 4096          # 8 x 4 B x 2^18 row blocks x 4096 = 2^35 bytes
    m = capi.V1DevMatrix(rows=rows, cols=cols, streams=32, precision=8, table_mode=0, coder=1,
                         payload_size=1 << 20, cursors=fake, lut=fake, payload=fake)
    st = L.matvans_v1_decompress_backward_noalloc(ctypes.byref(m), fake, fake, 1, rows, fake,
                                                  fake, None)
    assert st == 2 and "4 GiB" in capi.last_error()     # MATVANS_E_SHAPE
    st = L.matvans_decompress_backward(fake, fake, 1, rows, fake, fake, rows, cols, fake, 0, 8,
                                       fake, fake, fake, 1 << 20, fake, 1 << 40, None)
    assert st == 2 and "4 GiB" in capi.last_error()


def test_coder2_payload_must_hold_the_row_tables():
    """Coder 2 reads each row's table (streams + 2 u16 words) from the payload head by row
    index; every entry point refuses a payload_size shorter than those tables before any
    device call (the pointers below are never dereferenced)."""
    L = matvans_amd.lib()
    fake = ctypes.c_void_p(256)
    rows, cols, S = 64, 1024, 32
    m = capi.V1DevMatrix(rows=rows, cols=cols, streams=S, precision=8, table_mode=0, coder=2,
                         payload_size=rows * (S + 2) - 1, cursors=fake, lut=fake, payload=fake)
    st = L.matvans_v1_decompress_and_multiply(ctypes.byref(m), fake, fake, 1, cols, None)
    assert st == 2 and "row tables" in capi.last_error()
    st = L.matvans_v1_decompress_backward(ctypes.byref(m), fake, fake, 1, rows, fake, fake,
                                          fake, 1 << 30, None)
    assert st == 2 and "row tables" in capi.last_error()
    st = L.matvans_v1_decompress_backward_noalloc(ctypes.byref(m), fake, fake, 1, rows, fake,
                                                  fake, None)
    assert st == 2 and "row tables" in capi.last_error()
