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
