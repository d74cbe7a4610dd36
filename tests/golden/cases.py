"""Seeded inputs behind the golden fixtures (shared by make_golden.py and the tests).

Every array is regenerated from numpy's PCG64 streams; manifest.json records a
SHA-256 of each generated input so a change in the generator is detected
instead of silently producing a mismatch.
"""
import hashlib

import numpy as np

# name -> (n, dim, nlist, bits, seed, prune_alpha, empty_list, nq)
INDEX_CASES = {
    "d128b4": (3000, 128, 17, 4, 42, 0.0, True, 12),
    "d100b1": (2000, 100, 12, 1, 99, 0.0, False, 12),
    "d128b2": (2000, 128, 16, 2, 3, 0.01, False, 12),
    "d64b8": (1200, 64, 10, 8, 11, 0.0, False, 12),
    "d64b9": (1500, 64, 8, 9, 7, 0.0, False, 12),
    "d768b4": (700, 768, 6, 4, 5, 0.0, False, 8),
}

ROTATE_CASES = [(64, 7), (100, 99), (128, 42), (768, 5), (960, 0x5EED1234), (1024, 1)]
ENCODE_PADDED = 128
ENCODE_ROWS = 24
ACCUM_BLOCKS = [3, 16, 64, 200]


def mixture(name):
    """Gaussian mixture: centers ~ N(0, 16 I), points = center + N(0, I)."""
    n, dim, nlist, _bits, seed, _alpha, empty, nq = INDEX_CASES[name]
    rng = np.random.default_rng(seed)
    centers = (rng.standard_normal((nlist, dim)) * 4.0).astype(np.float32)
    used = nlist - 1 if empty else nlist
    labels = rng.integers(0, used, size=n).astype(np.uint32)
    data = (centers[labels] + rng.standard_normal((n, dim)).astype(np.float32)).astype(np.float32)
    # One row sits exactly on its centroid (zero residual path).
    data[5] = centers[labels[5]]
    qlab = rng.integers(0, used, size=nq)
    queries = (centers[qlab] + rng.standard_normal((nq, dim)).astype(np.float32)).astype(np.float32)
    return data, centers, labels, queries


def search_grid(name):
    """(k, nprobe, disable_prune) triples run against each fixture index."""
    nlist = INDEX_CASES[name][2]
    grid = [(10, 1, 0), (10, 4, 0), (10, nlist, 0), (10, nlist, 1)]
    if name == "d128b4":
        grid.append((100, 8, 0))
        grid.append((128, nlist, 1))
    return grid


def rotate_input(dim):
    rng = np.random.default_rng(1000 + dim)
    return rng.standard_normal((5, dim)).astype(np.float32)


def encode_input():
    rng = np.random.default_rng(2024)
    u = rng.standard_normal((ENCODE_ROWS, ENCODE_PADDED))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    return u.astype(np.float32)


def accum_input(blocks):
    rng = np.random.default_rng(20260920 + blocks)
    payload = rng.integers(0, 256, size=(2, blocks, 32), dtype=np.uint8)
    lut = rng.integers(0, 256, size=(blocks, 64), dtype=np.uint8)
    return payload, lut


def route_input():
    rng = np.random.default_rng(77)
    return (rng.standard_normal((10, 128)).astype(np.float32),
            rng.standard_normal((300, 128)).astype(np.float32))


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
