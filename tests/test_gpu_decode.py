"""GPU parity of the fused decode-matvec against the oracle (bit-exact int32).

Every launch goes through the C ABI of libmatvans.so; the oracle (oracle/) only checks
the results.  Cases: the golden files written by the reference generator, the reference
chain hash, small ragged shapes, 1..8 vectors, and the full LLM shapes."""
import ctypes
import os
import sys

import numpy as np
import pytest

import matvans_amd as mv
from matvans_amd.capi import MatvansError
from oracle import oracle as O

from .helpers import gaussian_int8, gaussian_vec, laplace_int8, uniform_vec

pytestmark = pytest.mark.gpu

FIXTURES = ["compressed_matrices_512_5bit_count2.bin", "compressed_matrices_1024_5bit_count1.bin"]


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def oracle_rec(cm):
    return O.Record(cm.rows, cm.cols, cm.grid_spacing, cm.cursors, cm.payload_size,
                    cm.min_value, cm.G, cm.cdf_data, cm.ppf_data, cm.payload)


def gpu_matvec(cm, vecs):
    import torch
    v = torch.from_numpy(np.ascontiguousarray(vecs)).cuda()
    out = cm.multiply(v)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("fx", FIXTURES)
def test_golden_files(cuda, golden_dir, fx):
    hdr, vec, mats = mv.read_bin(os.path.join(golden_dir, fx))
    for k, cm in enumerate(mats):
        note(f"{fx} matrix {k} rows={cm.rows} cols={cm.cols}")
        cm.to_device()
        dense = O.c_decode(oracle_rec(cm))
        vecs = np.stack([vec[:cm.cols], uniform_vec(cm.cols, seed=k),
                         gaussian_vec(cm.cols, seed=k)])
        got = gpu_matvec(cm, vecs)
        for i in range(3):
            assert np.array_equal(got[i], O.matvec_i32(dense, vecs[i]))
            assert np.array_equal(got[i], O.c_matvec(oracle_rec(cm), vecs[i]))


def test_golden_chain_hash(cuda, golden_dir):
    import torch
    hdr, vec, mats = mv.read_bin(os.path.join(golden_dir, FIXTURES[0]))
    note("chain over the 512 golden file")
    for m in mats:
        m.to_device()
    out = mv.run_chain(mats, torch.from_numpy(vec).cuda())
    torch.cuda.synchronize()
    assert mv.hash_int8_array(out.cpu().numpy()) == hdr["result_hash"]
    # per-step: int8 output and delta equal the fp32 oracle epilogue bit for bit
    v = vec
    d_vec = torch.from_numpy(vec).cuda()
    for m in mats:
        r8 = torch.empty(m.rows, dtype=torch.int8, device="cuda")
        r32 = torch.empty(m.rows, dtype=torch.int32, device="cuda")
        delta = m.decompressAndMult(r8, r32, d_vec)
        ref8, refd = O.requant_f32(O.matvec_i32(O.c_decode(oracle_rec(m)), v))
        assert np.float32(delta) == refd
        assert np.array_equal(r8.cpu().numpy(), ref8)
        v, d_vec = ref8, r8


SMALL = [  # rows, cols, generator, parameter
    (1, 512, "g", 6.0), (2, 512, "g", 6.0), (3, 1024, "g", 2.0), (7, 512, "g", 0.3),
    (8, 1536, "l", 4.0), (9, 2048, "g", 16.0), (15, 512, "g", 60.0), (64, 4096, "l", 4.0),
    (129, 3072, "g", 6.0), (1000, 1024, "g", 6.0), (333, 8192, "l", 4.0)]


@pytest.mark.parametrize("rows,cols,kind,p", SMALL)
def test_shapes_and_vectors(cuda, rows, cols, kind, p):
    note(f"shape {rows}x{cols} {kind}{p}")
    maker = gaussian_int8 if kind == "g" else laplace_int8
    m, _ = maker(rows, cols, p, seed=rows * 7 + cols)
    cm = mv.CompressedMatrix.encode(m).to_device()
    for n_vec in (1, 3, 8):
        vecs = np.stack([uniform_vec(cols, seed=100 + i) for i in range(n_vec)])
        got = gpu_matvec(cm, vecs)
        want = (m.astype(np.int64) @ vecs.astype(np.int64).T).T.astype(np.int32)
        assert np.array_equal(got, want), (n_vec, np.argwhere(got != want)[:5])


@pytest.mark.parametrize("val", [-127, 0, 127])
def test_constant_matrix(cuda, val):
    note(f"constant {val}")
    m = np.full((6, 1024), val, np.int8)
    cm = mv.CompressedMatrix.encode(m).to_device()
    v = np.full(1024, -127, np.int8)
    assert np.array_equal(gpu_matvec(cm, v[None])[0], O.matvec_i32(m, v))


LLM = [(4096, 4096, "l", 4.0), (14336, 4096, "l", 4.0), (4096, 14336, "l", 4.0),
       (4096, 14336, "g", 6.0)]


@pytest.mark.parametrize("rows,cols,kind,p", LLM)
def test_llm_shapes(cuda, rows, cols, kind, p):
    note(f"LLM shape {rows}x{cols} {kind}{p}")
    maker = gaussian_int8 if kind == "g" else laplace_int8
    m, _ = maker(rows, cols, p, seed=rows + cols)
    cm = mv.CompressedMatrix.encode(m, n_threads=16).to_device()
    vecs = np.stack([gaussian_vec(cols, seed=3), uniform_vec(cols, seed=4)])
    got = gpu_matvec(cm, vecs)
    want = (m.astype(np.int64) @ vecs.astype(np.int64).T).T.astype(np.int32)
    assert np.array_equal(got, want)
    assert np.array_equal(got[0], O.c_matvec(oracle_rec(cm), vecs[0], 16))


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_decode_variants_agree(cuda, golden_dir, variant):
    """Every decode kernel variant gives the oracle's bits (golden file + a ragged shape)."""
    note(f"decode variant {variant}")
    lib = mv.lib()
    assert lib.matvans_set_decode_variant(variant) == 0
    try:
        hdr, vec, mats = mv.read_bin(os.path.join(golden_dir, FIXTURES[1]))
        cm = mats[0].to_device()
        v = vec[:cm.cols]
        assert np.array_equal(gpu_matvec(cm, v[None])[0], O.c_matvec(oracle_rec(cm), v))
        m, _ = laplace_int8(77, 3072, 4.0, seed=variant)
        cm = mv.CompressedMatrix.encode(m).to_device()
        vecs = np.stack([uniform_vec(3072, seed=i) for i in range(5)])
        want = (m.astype(np.int64) @ vecs.astype(np.int64).T).T.astype(np.int32)
        assert np.array_equal(gpu_matvec(cm, vecs), want)
    finally:
        lib.matvans_set_decode_variant(-1)


def test_reference_signature_entry(cuda):
    import torch
    note("decmpressAndMultiply argument list")
    m, _ = gaussian_int8(40, 1024, 6.0, seed=2)
    cm = mv.CompressedMatrix.encode(m).to_device()
    v = uniform_vec(1024, seed=9)
    dv = torch.from_numpy(v).cuda()
    dst = torch.empty(40, dtype=torch.int32, device="cuda")
    mv.decmpressAndMultiply(dst, dv, cm.rows, cm.cols, 1.0, cm.d_cursors, cm.min_value, cm.G,
                            cm.d_cdf_data, cm.d_ppf_data, cm.d_payload, cm.payload_size)
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy(), O.matvec_i32(m, v))


def test_side_stream(cuda):
    import torch
    note("launch on a side stream")
    m, _ = gaussian_int8(100, 2048, 6.0, seed=4)
    cm = mv.CompressedMatrix.encode(m).to_device()
    v = uniform_vec(2048, seed=5)
    dv = torch.from_numpy(v).cuda()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        out = cm.multiply(dv, stream=s)
    s.synchronize()
    assert np.array_equal(out.cpu().numpy(), O.matvec_i32(m, v))


def test_argument_errors_do_not_launch(cuda):
    import torch
    note("argument validation")
    m, _ = gaussian_int8(8, 512, 6.0)
    cm = mv.CompressedMatrix.encode(m).to_device()
    dv = torch.zeros(1024, dtype=torch.int8, device="cuda")
    dst = torch.empty(8, dtype=torch.int32, device="cuda")
    with pytest.raises(MatvansError) as e:   # payload not 16-byte aligned
        mv.capi.call("matvans_decompress_and_multiply", ctypes.c_void_p(dst.data_ptr()),
                     ctypes.c_void_p(dv.data_ptr()), 8, 512, 1.0,
                     ctypes.c_void_p(cm.d_cursors.data_ptr()), cm.min_value, cm.G,
                     ctypes.c_void_p(cm.d_cdf_data.data_ptr()),
                     ctypes.c_void_p(cm.d_ppf_data.data_ptr()),
                     ctypes.c_void_p(cm.d_payload.data_ptr() + 2), cm.payload_size, None)
    assert e.value.status == 3
    with pytest.raises(MatvansError) as e:   # cols not a multiple of 512
        mv.decmpressAndMultiply(dst, dv, 8, 500, 1.0, cm.d_cursors, cm.min_value, cm.G,
                                cm.d_cdf_data, cm.d_ppf_data, cm.d_payload, cm.payload_size)
    assert e.value.status == 2
    v9 = torch.zeros((9, 512), dtype=torch.int8, device="cuda")
    with pytest.raises(ValueError, match="9 vectors"):   # the binding's check
        cm.multiply(v9)
    out9 = torch.empty((9, 8), dtype=torch.int32, device="cuda")
    with pytest.raises(MatvansError) as e:   # the C ABI's own check: 9 vectors
        mv.capi.call("matvans_decompress_and_multiply_nvec", ctypes.c_void_p(out9.data_ptr()),
                     ctypes.c_void_p(v9.data_ptr()), 9, 512, 8, 512,
                     ctypes.c_void_p(cm.d_cursors.data_ptr()), cm.min_value, cm.G,
                     ctypes.c_void_p(cm.d_cdf_data.data_ptr()),
                     ctypes.c_void_p(cm.d_ppf_data.data_ptr()),
                     ctypes.c_void_p(cm.d_payload.data_ptr()), cm.payload_size, None)
    assert e.value.status == 1


def test_library_upload_matches_python_upload(cuda, golden_dir):
    """matvans_matrix_upload / matvans_v1_matrix_upload (library-owned device copies)
    give the same decode as the Python-side upload."""
    import ctypes
    import torch
    from matvans_amd.capi import DevMatrix, HostMatrix, V1DevMatrix, call, lib
    from matvans_amd.mvx import CompressedMatrixV1
    hdr, vec, mats = mv.read_bin(os.path.join(golden_dir, FIXTURES[0]))
    m = mats[0]
    keep = [m.cursors, m.cdf_data, m.ppf_data, m.payload]
    h = HostMatrix(m.rows, m.cols, 1.0, m.min_value, m.G, 0, m.payload_size,
                   m.cursors.ctypes.data, m.cdf_data.ctypes.data, m.ppf_data.ctypes.data,
                   m.payload.ctypes.data)
    up, view = ctypes.c_void_p(), DevMatrix()
    call("matvans_matrix_upload", ctypes.byref(h), ctypes.byref(up), ctypes.byref(view))
    try:
        v = torch.from_numpy(vec[:m.cols].copy()).cuda()
        out = torch.empty(m.rows, dtype=torch.int32, device="cuda")
        call("matvans_decompress_and_multiply", ctypes.c_void_p(out.data_ptr()),
             ctypes.c_void_p(v.data_ptr()), view.rows, view.cols, view.grid_spacing,
             ctypes.c_void_p(view.cursors), view.min_value, view.G, ctypes.c_void_p(view.cdf),
             ctypes.c_void_p(view.ppf), ctypes.c_void_p(view.payload), view.payload_size, None)
        torch.cuda.synchronize()
        ref = m.to_device("cuda").multiply(v).cpu().numpy()
        assert np.array_equal(out.cpu().numpy(), ref)
    finally:
        lib().matvans_upload_free(up)
    del keep
    w = np.clip(np.rint(np.random.default_rng(3).normal(0, 6, (100, 2048))), -127, 127).astype(np.int8)
    c = CompressedMatrixV1.encode(w, 64, 8, 1, coder=1)
    keep = []
    hv = c.host_struct(keep)
    up, view = ctypes.c_void_p(), V1DevMatrix()
    call("matvans_v1_matrix_upload", ctypes.byref(hv), ctypes.byref(up), ctypes.byref(view))
    try:
        x = torch.from_numpy(np.arange(2048, dtype=np.int64).astype(np.int8)).cuda()
        out = torch.empty(100, dtype=torch.int32, device="cuda")
        call("matvans_v1_decompress_and_multiply", ctypes.byref(view),
             ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(x.data_ptr()), 1, 2048, None)
        torch.cuda.synchronize()
        want = (w.astype(np.int64) @ np.arange(2048).astype(np.int8).astype(np.int64)).astype(np.int32)
        assert np.array_equal(out.cpu().numpy(), want)
    finally:
        lib().matvans_upload_free(up)
