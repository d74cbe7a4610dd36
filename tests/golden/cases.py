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

# ME8Q: name -> (n, dim, nlist, avg_bits, seed, prune_alpha, plan_spec, pass_variance,
#                drop_energy, empty_list, nq)
ME8Q_CASES = {
    "m400a4": (1000, 400, 9, 4.0, 0, 0.0, "", False, True, False, 10),
    "m256sp": (1200, 256, 8, 4.0, 7, 0.01, "64:8,64:3,64:1,64:0", False, True, True, 10),
    "m128u4": (1000, 128, 6, 4.0, 3, 0.0, "128:4", False, True, False, 10),
    "m512v3": (900, 512, 7, 3.0, 11, 0.0, "", True, False, False, 10),
    "m256hi": (900, 256, 5, 4.0, 5, 0.0, "64:13,64:6,128:2", False, True, False, 8),
    "m64drop": (300, 64, 4, 2.0, 9, 0.0, "", False, True, False, 6),
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


def me8q_mixture(name):
    """Anisotropic mixture with descending per-dimension variance (the PCA
    domain ME8Q consumes): lambda_i = i^-1.2, centers ~ N(0, 9 diag(lambda)),
    points = center + N(0, diag(lambda)).  Returns data, centers, labels,
    queries and the generating variance."""
    n, dim, nlist, _b, seed, _a, _spec, _pv, _de, empty, nq = ME8Q_CASES[name]
    rng = np.random.default_rng(1000 + seed)
    lam = (np.arange(1, dim + 1, dtype=np.float64) ** -1.2).astype(np.float32)
    sd = np.sqrt(lam).astype(np.float32)
    centers = (rng.standard_normal((nlist, dim)).astype(np.float32) * (3.0 * sd)).astype(np.float32)
    used = nlist - 1 if empty else nlist
    labels = rng.integers(0, used, size=n).astype(np.uint32)
    data = (centers[labels] + rng.standard_normal((n, dim)).astype(np.float32) * sd).astype(np.float32)
    data[3] = centers[labels[3]]
    qlab = rng.integers(0, used, size=nq)
    queries = (centers[qlab] + rng.standard_normal((nq, dim)).astype(np.float32) * sd).astype(np.float32)
    return data, centers, labels, queries, lam


def me8q_search_grid(name):
    nlist = ME8Q_CASES[name][2]
    return [(10, 1), (10, 3), (10, nlist), (100, nlist), (7, nlist)]


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


# ---------------------------------------------------------------------------
# Configuration-shaped fixtures (make_golden.py --only flush,c1,c3): the index
# image is kept as SHA-256 digests, the search results in full.

# Survivor flush sizes (E8Q_FLUSH_EVERY) run against these fixture indexes:
# name -> [(k, nprobe, disable_prune)]
FLUSH_SIZES = [1, 7, 256, 1024]


def flush_grid(name):
    nlist = INDEX_CASES[name][2]
    return [(10, nlist, 0), (100, 8 if nlist > 8 else nlist, 0)]


FLUSH_INDEXES = ["d128b4", "d768b4"]

# The CPU-reference E8Q configuration at full size: N = 100k, D = 128 i.i.d.
# N(0, 1) (numpy PCG64 seed 42), then 1000 queries from the same stream,
# 256 centroids = distinct random base rows from the same stream,
# cluster_ids = nearest centroid (fp64, kept in the fixture); b = 4,
# rotator seed 42, k = 10, nprobe = 16.
C1 = dict(n=100_000, dim=128, nlist=256, nq=1000, bits=4, seed=42, k=10, nprobe=16)


def c1_iid():
    rng = np.random.default_rng(C1["seed"])
    base = rng.standard_normal((C1["n"], C1["dim"]), dtype=np.float32)
    queries = rng.standard_normal((C1["nq"], C1["dim"]), dtype=np.float32)
    pick = rng.choice(C1["n"], C1["nlist"], replace=False)
    return base, np.ascontiguousarray(base[pick]), queries


def nearest_centroid(x, cent, chunk=4096):
    """argmin_c ||x - c||^2 in fp64 (the fixture keeps the result)."""
    c64 = cent.astype(np.float64)
    cn = (c64 * c64).sum(1)
    out = np.empty(x.shape[0], np.int64)
    for s in range(0, x.shape[0], chunk):
        xb = x[s:s + chunk].astype(np.float64)
        out[s:s + chunk] = np.argmin(cn[None, :] - 2.0 * (xb @ c64.T), axis=1)
    return out


# The mixed-rate ME8Q configuration's shape: D = 960 with the power-law
# spectrum lambda_i = i^-1.2 (the PCA domain's descending variance, passed as
# `variance`), 256 lists = seeded random base rows, cluster_ids = nearest
# (kept in the fixture), the SAQ DP plan at avg_bits 4, k = 100 / 10,
# nprobe 32 / 64.
C3 = dict(n=20_000, dim=960, nlist=256, nq=32, avg_bits=4.0, seed=0, components=512)


def c3_shape():
    rng = np.random.default_rng(960)
    dim = C3["dim"]
    lam = (np.arange(1, dim + 1, dtype=np.float64) ** -1.2).astype(np.float32)
    sd = np.sqrt(lam).astype(np.float32)
    comp = (rng.standard_normal((C3["components"], dim)).astype(np.float32) * (2.0 * sd)).astype(np.float32)
    lab = rng.integers(0, C3["components"], size=C3["n"])
    data = (comp[lab] + rng.standard_normal((C3["n"], dim)).astype(np.float32) * sd).astype(np.float32)
    qlab = rng.integers(0, C3["components"], size=C3["nq"])
    queries = (comp[qlab] + rng.standard_normal((C3["nq"], dim)).astype(np.float32) * sd).astype(np.float32)
    pick = rng.choice(C3["n"], C3["nlist"], replace=False)
    return data, np.ascontiguousarray(data[pick]), queries, lam


def c3_grid():
    return [(100, 32), (100, 64), (10, 32)]


def e8ax_digests(ix):
    """SHA-256 of each array section of an E8AX image (tests/e8ax.py read())."""
    return {key: digest(np.asarray(ix[key])) for key in ("ids", "signs", "mean_rot", "centroid_rot", "centroid_n2",
                                                          "cluster_begin", "arena_begin", "arena", "ref_arena_begin",
                                                          "ref_arena")}


# Diagnostics (Estimate, ProbeBoundW, ExhaustiveTopK, SearchIVF) over the
# fixture indexes: (query, id) pairs, six random ids per query plus the row
# that sits exactly on its centroid (row 5 of an E8Q mixture, row 3 of an
# ME8Q one) for the first and the last query.
def diag_pairs(name):
    if name in INDEX_CASES:
        n, nq, seed, zero = INDEX_CASES[name][0], INDEX_CASES[name][7], INDEX_CASES[name][4], 5
    else:
        n, nq, seed, zero = ME8Q_CASES[name][0], ME8Q_CASES[name][10], ME8Q_CASES[name][4], 3
    rng = np.random.default_rng(7000 + seed)
    pq = np.repeat(np.arange(nq), 6)
    pid = rng.integers(0, n, size=pq.size)
    pq = np.concatenate([pq, [0, nq - 1]])
    pid = np.concatenate([pid, [zero, zero]])
    return pq.astype(np.uint32), pid.astype(np.uint32)


def diag_k(name):
    """k of E8Q ExhaustiveTopK (ME8Q: the k of diag_ivf_grid)."""
    return [10, 64]


def diag_ivf_grid(name):
    """(k, nprobe) of ME8Q SearchIVF."""
    nlist = ME8Q_CASES[name][2]
    grid = me8q_search_grid(name)
    return grid + [(400, nlist)] if name == "m64drop" else grid
