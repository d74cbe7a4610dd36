"""e8qlib.E8qIndex (host C++ over the device library) as a drop-in for the
reference Python API: build/search/save/load round trips, E8AX files that
are byte-identical to the reference's, and search results equal to the
reference's on its own files."""
import os

import numpy as np
import pytest

import cases
import e8qlib

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    # Same shape as the reference's tests/python/test_indexes.py:14-25.
    rng = np.random.default_rng(42)
    dim, n, nq, nl = 64, 400, 16, 8
    scale = np.power(0.97, np.arange(dim)).astype(np.float32)
    base = (rng.standard_normal((n, dim)) * scale).astype(np.float32)
    queries = (rng.standard_normal((nq, dim)) * scale).astype(np.float32)
    cid = (np.arange(n) % nl).astype(np.uint32)
    cent = np.stack([base[cid == c].mean(axis=0) for c in range(nl)]).astype(np.float32)
    return base, queries, cent, cid


def test_build_search_roundtrip(data, tmp_path):
    base, queries, cent, cid = data
    idx = e8qlib.E8qIndex(dim=64, nbits=4, seed=1)
    idx.build(base, cent, cid, num_threads=2)
    assert idx.is_built and idx.n == 400 and idx.num_clusters == 8
    ids, dists = idx.search(queries, k=10, nprobe=8, num_threads=2)
    assert ids.shape == (16, 10) and dists.shape == (16, 10)
    assert np.all(ids < 400)
    assert np.all(np.diff(dists, axis=1) >= 0)
    path = tmp_path / "e8q.bin"
    idx.save(str(path))
    loaded = e8qlib.E8qIndex.load(str(path))
    ids2, d2 = loaded.search(queries, k=10, nprobe=8)
    np.testing.assert_array_equal(ids, ids2)
    np.testing.assert_array_equal(dists, d2)


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_saved_file_matches_reference_bytes(name, tmp_path):
    n, dim, nlist, bits, seed, alpha, _e, _nq = cases.INDEX_CASES[name]
    data, cent, lab, _q = cases.mixture(name)
    idx = e8qlib.E8qIndex(dim, bits, seed)
    print(f"build {name}", flush=True)
    idx.build(data, cent, lab, prune_alpha=alpha)
    out = tmp_path / f"{name}.e8ax"
    idx.save(str(out))
    assert out.read_bytes() == open(os.path.join(G, f"{name}.e8ax"), "rb").read()


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_reference_file_search_matches(name):
    ref = np.load(os.path.join(G, f"{name}_search.npz"))
    idx = e8qlib.E8qIndex.load(os.path.join(G, f"{name}.e8ax"))
    q = ref["queries"]
    for (k, nprobe, dis) in cases.search_grid(name):
        tag = f"k{k}_p{nprobe}_d{dis}"
        print(f"search {name} {tag}", flush=True)
        idx.set_disable_prune(bool(dis))
        ids, d, st = idx.search(q, k, nprobe, return_stats=True)
        np.testing.assert_array_equal(ids, ref[f"{tag}_ids"], err_msg=tag)
        np.testing.assert_array_equal(d, ref[f"{tag}_dists"], err_msg=tag)
        assert st["candidates"] == int(ref[f"{tag}_stats"][:, 0].sum())
        assert st["survivors"] == int(ref[f"{tag}_stats"][:, 1].sum())


def test_search_errors(data):
    base, queries, cent, cid = data
    idx = e8qlib.E8qIndex(64, 4)
    idx.build(base, cent, cid)
    with pytest.raises(ValueError, match="k must be <= 128"):
        idx.search(queries, 129, 4)
    with pytest.raises(ValueError, match="k must be positive"):
        idx.search(queries, 0, 4)
    with pytest.raises(ValueError, match="nprobe must be positive"):
        idx.search(queries, 10, 0)
    with pytest.raises(ValueError, match="query dimension does not match"):
        idx.search(queries[:, :32], 10, 4)
    with pytest.raises(ValueError, match="cluster id out of range"):
        e8qlib.E8qIndex(64, 4).build(base, cent, np.full(400, 8, np.uint32))
    # nprobe beyond nlist clamps; empty query set returns empty arrays.
    a, _ = idx.search(queries, 10, 1000)
    b, _ = idx.search(queries, 10, 8)
    np.testing.assert_array_equal(a, b)
    e, ed = idx.search(queries[:0], 10, 4)
    assert e.shape == (0, 10) and ed.shape == (0, 10)


def test_exhaustive_against_full_probe_search(data):
    # The reference's own check (e8q_test.cpp:281-373): the unpruned search
    # over every list returns ExhaustiveTopK's ids, up to a swap at the k-th
    # place within 1.0001 d_k + 1e-4 (their list distances are summed in
    # different orders).
    base, queries, cent, cid = data
    idx = e8qlib.E8qIndex(64, 4)
    idx.build(base, cent, cid)
    idx.set_disable_prune(True)
    ids, d = idx.search(queries, 10, 8)
    for i in range(queries.shape[0]):
        ex = idx.exhaustive_topk(queries[i], 10)
        assert len(ex) == 10
        exd = np.array([e[0] for e in ex], np.float32)
        assert np.all(np.diff(exd) >= 0)
        lim = 1.0001 * exd[-1] + 1e-4
        for a, b in zip(ids[i], d[i]):
            if int(a) not in {e[1] for e in ex}:
                assert b <= lim
