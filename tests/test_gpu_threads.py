"""Concurrent use from several host threads, each with its own stream and its own
matrices: the .bin v0 composite (decode-matvec, device absmax, requant), the quad-coder
forward and a batched plan, repeated, every result bit-exact against the oracle.  The
library's launch path keeps no shared mutable host state (error strings are per thread,
device caches are read-only after their first fill) and ctypes releases the GIL, so the
threads' calls overlap."""
import sys
import threading

import numpy as np
import pytest

import matvans_amd as mv
from matvans_amd.batch import Batch
from matvans_amd.mvx import CompressedMatrixV1
from oracle import oracle as O

from .helpers import gaussian_int8, uniform_vec

pytestmark = pytest.mark.gpu

N_THREADS, ITERS = 4, 20


def note(msg):
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def test_threads_on_their_own_streams(cuda):
    import torch
    note(f"{N_THREADS} host threads, own streams: v0 composite, quad forward, batch")
    cases = []
    for i in range(N_THREADS):
        rows, cols = 96 + 40 * i, 2048
        m, _ = gaussian_int8(rows, cols, 3.0 + i, seed=40 + i)
        v = uniform_vec(cols, seed=60 + i)
        cm = mv.CompressedMatrix.encode(m).to_device()
        q = CompressedMatrixV1.encode(m, 32, 8, 0, coder=1).to_device()
        dv = torch.from_numpy(v).cuda()
        want32 = O.matvec_i32(m, v)
        want8, want_delta = O.requant_f32(want32)
        cases.append(dict(
            cm=cm, q=q, dv=dv, want32=want32, want8=want8, want_delta=want_delta,
            r32=torch.empty((ITERS, rows), dtype=torch.int32, device="cuda"),
            r8=torch.empty((ITERS, rows), dtype=torch.int8, device="cuda"),
            delta=torch.empty(ITERS, dtype=torch.float32, device="cuda"),
            q32=torch.empty((ITERS, rows), dtype=torch.int32, device="cuda"),
            batch=Batch([(q, dv), (cm, dv)]),
            b32=torch.empty((ITERS, 2, rows), dtype=torch.int32, device="cuda")))
    torch.cuda.synchronize()
    errors = []
    gate = threading.Barrier(N_THREADS)

    def work(c):
        try:
            s = torch.cuda.Stream()
            gate.wait()
            with torch.cuda.stream(s):
                for it in range(ITERS):
                    c["cm"].decompress_and_mult(c["r8"][it], c["r32"][it], c["dv"],
                                                c["delta"][it:it + 1], stream=s)
                    c["q"].multiply(c["dv"], out=c["q32"][it:it + 1], stream=s)
                    c["batch"].launch(stream=s)
                    for k in range(2):
                        c["b32"][it, k].copy_(c["batch"].outputs[k][0])
            s.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    ts = [threading.Thread(target=work, args=(c,)) for c in cases]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    assert not errors, errors
    for i, c in enumerate(cases):
        for it in range(ITERS):
            assert np.array_equal(c["r32"][it].cpu().numpy(), c["want32"]), (i, it)
            assert np.array_equal(c["r8"][it].cpu().numpy(), c["want8"]), (i, it)
            assert c["delta"][it].item() == c["want_delta"], (i, it)
            assert np.array_equal(c["q32"][it].cpu().numpy(), c["want32"]), (i, it)
            for k in range(2):
                assert np.array_equal(c["b32"][it, k].cpu().numpy(), c["want32"]), (i, it, k)
