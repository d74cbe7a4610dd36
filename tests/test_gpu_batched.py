"""The batched search form (E8Q_MODE_BATCHED) against the plain form.

The batched form fixes each query's prune threshold at its top-k k-th
distance after the first probe (the reference's running threshold at that
point) and tests every later (query, cluster) pair against it; the bound only
rises as the threshold falls, so its candidates contain the plain search's.
Hence, exactly (float compare, no tolerance):
  * prune off: ids and distances equal the plain search's;
  * prune on: every rank's distance is <= the plain search's, and the first
    probe's candidates and the candidate/batch counts are the same.
"""
import os

import numpy as np
import pytest

import capi
import cases
import e8ax
import oracle_py as orc
from tables import empty_arrays, layout, tables

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


def _z(k, alpha=1.0):
    return orc.lib().orc_ortho_quantile(alpha, max(4 * k, 32))


def _mixture_index(ctx, n, dim, nlist, bits, seed, nq, sizes=None):
    rng = np.random.default_rng(seed)
    cent = (rng.standard_normal((nlist, dim)) * 2.0).astype(np.float32)
    if sizes is None:
        lab = rng.integers(0, nlist, n).astype(np.uint32)
    else:
        lab = np.repeat(np.arange(nlist, dtype=np.uint32), sizes)
        n = lab.size
    data = (cent[lab] + rng.standard_normal((n, dim)).astype(np.float32)).astype(np.float32)
    q = (cent[rng.integers(0, nlist, nq)] + rng.standard_normal((nq, dim)).astype(np.float32)).astype(np.float32)
    lay = layout(lab, nlist, bits, dim)
    arr = empty_arrays(n, dim, nlist, bits, lay, orc.rotator_signs(dim, seed))
    ix = capi.DeviceIndex(ctx, arr, tables(bits, lay["padded"]))
    ix.encode(data, cent, lab, lay["slot"])
    return ix, q


def _check_never_worse(plain, batched):
    (pi, pd, _), (bi, bd, _) = plain, batched
    assert np.all(bd <= pd), "a batched rank is worse than the plain search's"
    # where the distances agree the ids agree too (ties broken by id)
    same = bd == pd
    assert np.array_equal(bi[same], pi[same])


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_prune_off_equals_plain(ctx, name):
    a = e8ax.read(os.path.join(G, f"{name}.e8ax"))
    ix = capi.DeviceIndex(ctx, a, tables(a["bits"], a["padded"]))
    ix.upload()
    q = np.load(os.path.join(G, f"{name}_search.npz"))["queries"]
    nlist = a["nlist"]
    for k, nprobe in ((10, nlist), (10, 3), (min(128, a["n"]), nlist)):
        print(f"batched prune off {name} k={k} nprobe={nprobe}", flush=True)
        pids, pd, ps = ix.search(q, k, nprobe, _z(k), disable_prune=True)
        bids, bd, bs = ix.search(q, k, nprobe, _z(k), disable_prune=True, mode=capi.MODE_BATCHED)
        np.testing.assert_array_equal(bids, pids)
        np.testing.assert_array_equal(bd, pd)
        assert bs[0] == ps[0]


@pytest.mark.parametrize("name", ["d128b4", "d128b2", "d64b8", "d768b4"])
def test_prune_on_never_worse(ctx, name):
    a = e8ax.read(os.path.join(G, f"{name}.e8ax"))
    ix = capi.DeviceIndex(ctx, a, tables(a["bits"], a["padded"]))
    ix.upload()
    q = np.load(os.path.join(G, f"{name}_search.npz"))["queries"]
    for k, nprobe in ((10, a["nlist"]), (10, 4)):
        print(f"batched prune on {name} k={k} nprobe={nprobe}", flush=True)
        plain = ix.search(q, k, nprobe, _z(k))
        batched = ix.search(q, k, nprobe, _z(k), mode=capi.MODE_BATCHED)
        _check_never_worse(plain, batched)
        assert batched[2][0] == plain[2][0]   # candidates


def test_mixture_never_worse_and_recall(ctx):
    ix, q = _mixture_index(ctx, 40000, 128, 64, 4, 21, 300)
    for k, nprobe in ((100, 16), (10, 8), (10, 64)):
        print(f"batched mixture k={k} nprobe={nprobe}", flush=True)
        plain = ix.search(q, k, nprobe, _z(k))
        batched = ix.search(q, k, nprobe, _z(k), mode=capi.MODE_BATCHED)
        _check_never_worse(plain, batched)
        exact = ix.search(q, k, nprobe, _z(k), disable_prune=True)
        # recall against the unpruned estimator top-k: at least the plain search's
        hit = lambda r: sum(len(set(r[0][i]) & set(exact[0][i])) for i in range(q.shape[0]))
        assert hit(batched) >= hit(plain)


def test_uneven_nprobe(ctx):
    ix, q = _mixture_index(ctx, 30000, 128, 48, 4, 5, 200)
    rng = np.random.default_rng(3)
    npq = rng.choice(np.array([1, 2, 5, 16, 48], np.uint32), size=q.shape[0])
    print("batched uneven nprobe", flush=True)
    plain = ix.search(q, 10, 48, _z(10), nprobe_per_query=npq)
    batched = ix.search(q, 10, 48, _z(10), nprobe_per_query=npq, mode=capi.MODE_BATCHED)
    _check_never_worse(plain, batched)
    # a query with a single probe is its first-probe pass: identical
    one = npq == 1
    np.testing.assert_array_equal(batched[0][one], plain[0][one])
    np.testing.assert_array_equal(batched[1][one], plain[1][one])


def test_skewed_clusters_and_overflow(ctx):
    # One list far larger than the others (several items per pair tile, and,
    # with pruning off, more candidates than a query keeps: the plain fallback).
    sizes = np.array([6000, 40, 900, 3, 1200, 0, 700, 350], np.int64)
    ix, q = _mixture_index(ctx, 0, 64, 8, 4, 8, 60, sizes=sizes)
    for dis in (False, True):
        print(f"batched skewed disable_prune={dis}", flush=True)
        plain = ix.search(q, 50, 8, _z(50), disable_prune=dis)
        batched = ix.search(q, 50, 8, _z(50), disable_prune=dis, mode=capi.MODE_BATCHED)
        if dis:
            np.testing.assert_array_equal(batched[0], plain[0])
            np.testing.assert_array_equal(batched[1], plain[1])
        else:
            _check_never_worse(plain, batched)
