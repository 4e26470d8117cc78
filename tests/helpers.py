"""Seeded synthetic inputs shared by the tests, the smoke check and the bench."""
import numpy as np

SEED = 20250310


def gaussian_int8(rows, cols, sigma, seed=SEED):
    """W = clip(rint(N(0, sigma^2)), -127, 127) as int8, plus the fp64 source."""
    rng = np.random.default_rng(seed)
    wf = rng.normal(0.0, sigma, size=(rows, cols))
    return np.clip(np.rint(wf), -127, 127).astype(np.int8), wf


def laplace_int8(rows, cols, b, seed=SEED):
    rng = np.random.default_rng(seed)
    wf = rng.laplace(0.0, b, size=(rows, cols))
    return np.clip(np.rint(wf), -127, 127).astype(np.int8), wf


def uniform_vec(n, seed=SEED + 1, lo=-127, hi=127):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=n).astype(np.int8)


def gaussian_vec(n, sigma=30.0, seed=SEED + 2):
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(rng.normal(0, sigma, size=n)), -127, 127).astype(np.int8)


def gauss_table(sigma):
    """int8 values of clip(rint(N(0, sigma^2)), -127, 127) indexed by a uniform 16-bit
    draw (inverse-CDF sampling at 2^-16 resolution; tails folded into +-127): fast
    generation of the large batched workloads."""
    from math import erf, sqrt
    k = np.arange(-127, 128)
    cdf_hi = np.array([0.5 * (1 + erf((v + 0.5) / (sigma * sqrt(2)))) for v in k])
    cdf_hi[-1] = 1.0
    edges = np.rint(cdf_hi * 65536).astype(np.int64)
    return np.repeat(k.astype(np.int8), np.diff(np.concatenate([[0], edges])))[:65536]


def sample_int8(rng, shape, sigma):
    return gauss_table(sigma)[rng.integers(0, 65536, size=shape, dtype=np.uint16)]


def act_vectors(rng, n, cols):
    return np.clip(np.rint(rng.normal(0, 30.0, size=(n, cols))), -127, 127).astype(np.int8)


# the decoder block: q, k, v, o, gate, up, down (rows x cols)
BLOCK_SHAPES = [("q", 4096, 4096), ("k", 1024, 4096), ("v", 1024, 4096), ("o", 4096, 4096),
                ("gate", 14336, 4096), ("up", 14336, 4096), ("down", 4096, 14336)]


def uneven_specs(n_mat=32, seed=20250312, scale=1.0):
    """The uneven batch: n_mat matrices with rows ~ U[256, 16384] (multiples of 32), cols ~
    U{512, 768, ..., 8192}, entry sigma from {2, 4, 8, 16}, row or stripe layout (even
    odds; sigma = 16 matrices always take the row layout, because the stripe layout's
    7-bit model holds at most 128 distinct values and their range is wider), 1..8 vectors
    each.  Returns [(rows, cols, sigma, layout, n_vec)]."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_mat):
        rows = max(32, int(rng.integers(256 // 32, 16384 // 32 + 1) * 32 * scale) // 32 * 32)
        cols = max(512, int(rng.integers(2, 33) * 256 * scale) // 256 * 256)
        sigma = float(rng.choice([2.0, 4.0, 8.0, 16.0]))
        layout = "rows" if rng.random() < 0.5 or sigma == 16.0 else "stripes"
        specs.append((rows, cols, sigma, layout, int(rng.integers(1, 9))))
    return specs


def pad_cols(m, multiple):
    """m with zero columns appended up to a multiple of ``multiple`` (the product with a
    vector zero-padded the same way is unchanged)."""
    cols = m.shape[1]
    pc = -(-cols // multiple) * multiple
    if pc == cols:
        return m
    out = np.zeros((m.shape[0], pc), np.int8)
    out[:, :cols] = m
    return out
