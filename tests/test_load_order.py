"""The library and torch share one HIP runtime whatever the import order.

libmatvans.so links libamdhip64.so.7; torch carries its own copy.  Loading the library
before torch used to map ROCm's runtime, and the later ``import torch`` mapped torch's
beside it (two HIP and two HSA runtimes in one process).  ``capi.lib()`` now maps torch's
runtime first.  Each case runs in a fresh interpreter, because the test process may
already hold torch."""
import os
import subprocess
import sys
import textwrap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(code):
    r = subprocess.run([sys.executable, "-c", textwrap.dedent(code)], cwd=ROOT,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.strip().splitlines()


def test_library_first_then_torch_maps_one_runtime():
    out = run("""
        import matvans_amd
        from matvans_amd import capi
        capi.lib()
        import torch
        print(len(capi.hip_runtimes_mapped()))
        print(capi.hip_runtimes_mapped()[0])
    """)
    assert out[0] == "1", out
    assert "torch" in out[1], out


def test_torch_first_then_library_maps_one_runtime():
    out = run("""
        import torch
        from matvans_amd import capi
        capi.lib()
        print(len(capi.hip_runtimes_mapped()))
    """)
    assert out[0] == "1", out


def test_graft_build_then_smoke_order_maps_one_runtime():
    """__graft_entry__ build() loads the library before smoke() imports torch."""
    out = run("""
        import __graft_entry__ as g
        import matvans_amd
        matvans_amd.lib()
        import torch
        from matvans_amd import capi
        print(len(capi.hip_runtimes_mapped()))
    """)
    assert out[0] == "1", out


def test_second_runtime_is_refused(tmp_path):
    """A process that already mapped ROCm's runtime (another HIP library loaded before
    torch) gets a clear error from lib() instead of two runtimes."""
    out = run("""
        import ctypes
        ctypes.CDLL("libamdhip64.so.7")
        from matvans_amd import capi
        try:
            capi.lib()
        except RuntimeError as e:
            print("refused" if "two HIP runtimes" in str(e) else "other: " + str(e))
        else:
            print("loaded", len(capi.hip_runtimes_mapped()))
    """)
    assert out[-1] == "refused", out
