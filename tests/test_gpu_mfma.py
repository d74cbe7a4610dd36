"""The MFMA form (E8Q_MODE_MFMA): the unpruned refine estimator over every
probed vector as a split-BF16 contraction, FP32 accumulation.

Its reference result is the oracle's unpruned estimator top-k (ScanAllTopK,
the top-k the reference's prune-off search returns).  Checked here:
  * distance error against the FP64 oracle (orc_dist64: fp64 products and
    sum of the same decoded levels and query):
        |d_mfma - d_64| <= TOL * max|d_64| over the query's results,
    TOL = 2e-5 -- the hi/lo split leaves 2^-17 relative per term and FP32
    accumulates 2P terms (the fp32 refine chain itself is within 1e-5);
  * the ranking: the fp64 distances of the returned ids, sorted, equal the
    fp64 distances of the oracle's top-k within the same bound (ids may
    differ only where two candidates are that close);
  * the code paths around the contraction: first-probe lists shorter than k
    (unbounded candidates and the second pass with longer lists), uneven
    nprobe, graph replay, and the bits <= 4 restriction.
"""
import os

import numpy as np
import pytest

import capi
import e8ax
import oracle_py as orc
from tables import tables
from test_gpu_batched import _mixture_index, _z

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu
MODE_MFMA = 2
TOL = 2e-5
UMAX = np.iinfo(np.uint32).max


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


def _fixture(ctx, name):
    a = e8ax.read(os.path.join(G, f"{name}.e8ax"))
    ix = capi.DeviceIndex(ctx, a, tables(a["bits"], a["padded"]))
    ix.upload()
    q = np.load(os.path.join(G, f"{name}_search.npz"))["queries"]
    return ix, orc.OracleIndex.from_e8ax(a), q


def _oracle_of(ix):
    a = dict(ix.a)
    a.update(ix.download())
    a.setdefault("alpha", 1.0)
    return orc.OracleIndex.from_e8ax(a)


def _check(oix, q, k, nprobe, res, label):
    ids, d, _ = res
    ref_ids, _ = oix.scan_all(q, k, nprobe)
    d64 = oix.dist64(q, nprobe, ids)
    r64 = oix.dist64(q, nprobe, ref_ids)
    worst = 0.0
    for i in range(q.shape[0]):
        ok, rok = ids[i] != UMAX, ref_ids[i] != UMAX
        assert ok.sum() == rok.sum(), (label, i)
        assert np.all(d[i][~ok] == np.inf)
        if not ok.any():
            continue
        assert np.all(np.isfinite(d64[i][ok])), (label, i, "an id outside the probed lists")
        scale = max(np.abs(r64[i][rok]).max(), np.abs(d64[i][ok]).max())
        err = np.abs(d[i][ok].astype(np.float64) - d64[i][ok]).max() / scale
        worst = max(worst, err)
        assert err <= TOL, (label, i, err)
        # same fp64 distance profile as the oracle's top-k
        np.testing.assert_allclose(np.sort(d64[i][ok]), np.sort(r64[i][rok]), rtol=0, atol=2 * TOL * scale)
        assert len(set(ids[i][ok].tolist())) == ok.sum()
        assert np.all(np.diff(d[i][ok]) >= 0)
    print(f"{label}: max |d_mfma - d_fp64| / max|d| = {worst:.3e}", flush=True)
    return worst


@pytest.mark.parametrize("name", ["d128b4", "d100b1", "d128b2", "d768b4"])
def test_fixture_indexes(ctx, name):
    ix, oix, q = _fixture(ctx, name)
    nl = oix.nlist
    for k, nprobe in ((10, 1), (10, 3), (10, nl), (min(100, oix.n), nl)):
        print(f"mfma {name} k={k} nprobe={nprobe}", flush=True)
        res = ix.search(q, k, nprobe, _z(k), mode=MODE_MFMA)
        _check(oix, q, k, nprobe, res, f"{name} k={k} nprobe={nprobe}")
    ix.close()


def test_distances_come_from_the_contraction(ctx):
    # Same ids as the fp32 refine (prune off, all lists), but the distances are
    # the contraction's: within TOL of them and not bit-equal throughout.
    ix, oix, q = _fixture(ctx, "d128b4")
    nl = oix.nlist
    ref = ix.search(q, 10, nl, _z(10), disable_prune=True)
    res = ix.search(q, 10, nl, _z(10), mode=MODE_MFMA)
    same = res[0] == ref[0]
    assert same.mean() > 0.9
    dm, dr = res[1][same].astype(np.float64), ref[1][same].astype(np.float64)
    assert np.all(np.abs(dm - dr) <= 2 * TOL * np.abs(dr).max())
    assert np.mean(dm != dr) > 0.1
    ix.close()


def test_rejects_more_than_four_bits(ctx):
    ix, _, q = _fixture(ctx, "d64b8")
    with pytest.raises(RuntimeError, match="bits <= 4"):
        ix.search(q, 10, 3, _z(10), mode=MODE_MFMA)
    ix.close()


def test_mixture_recall_not_below_plain(ctx):
    ix, q = _mixture_index(ctx, 30000, 128, 48, 4, 31, 200)
    oix = _oracle_of(ix)
    for nprobe in (4, 16):
        print(f"mfma mixture nprobe={nprobe}", flush=True)
        res = ix.search(q, 10, nprobe, _z(10), mode=MODE_MFMA)
        _check(oix, q, 10, nprobe, res, f"mixture nprobe={nprobe}")
        # the unpruned result is never worse than the pruned plain search (fp64 distances)
        plain = ix.search(q, 10, nprobe, _z(10))
        dm = np.sort(oix.dist64(q, nprobe, res[0]), axis=1)
        dp = np.sort(oix.dist64(q, nprobe, plain[0]), axis=1)
        assert np.all(dm <= dp + 2 * TOL * np.abs(dp).max())
    ix.close()


def test_uneven_nprobe_and_second_pass(ctx):
    # One list of 6000 vectors and several tiny ones: a query whose first list
    # holds fewer than k vectors has no bound, keeps every candidate (> 4096
    # of them) and is scored again with a list of its length.
    sizes = np.array([6000, 40, 900, 3, 1200, 0, 700, 350], np.int64)
    ix, q = _mixture_index(ctx, 0, 64, 8, 4, 8, 60, sizes=sizes)
    oix = _oracle_of(ix)
    rng = np.random.default_rng(4)
    npq = rng.choice(np.array([1, 2, 5, 8], np.uint32), size=q.shape[0])
    for k in (5, 50):
        print(f"mfma skewed k={k}", flush=True)
        res = ix.search(q, k, 8, _z(k), mode=MODE_MFMA)
        _check(oix, q, k, 8, res, f"skewed k={k}")
        if k == 50:   # one query at a time: some keep more than the 4096 a list holds
            surv = []
            for i in range(q.shape[0]):
                one = ix.search(q[i:i + 1], k, 8, _z(k), mode=MODE_MFMA)
                np.testing.assert_array_equal(one[0][0], res[0][i])
                np.testing.assert_array_equal(one[1][0], res[1][i])
                surv.append(one[2][1])
            assert max(surv) > 4096, "the second pass was not exercised"
        resq = ix.search(q, k, 8, _z(k), mode=MODE_MFMA, nprobe_per_query=npq)
        for p in (1, 2, 5, 8):
            sel = npq == p
            if sel.any():
                _check(oix, q[sel], k, int(p), (resq[0][sel], resq[1][sel], None), f"skewed k={k} nprobe={p}")
    ix.close()


def test_graph_replay(ctx):
    ix, q = _mixture_index(ctx, 20000, 128, 32, 4, 12, 150)
    ctx.set_graphs(False)
    rng = np.random.default_rng(2)
    qs = [q] + [(q + 0.05 * rng.standard_normal(q.shape)).astype(np.float32) for _ in range(2)]
    eager = [ix.search(x, 10, 8, _z(10), mode=MODE_MFMA) for x in qs]
    ctx.set_graphs(True)
    try:
        for rep in range(3):
            for i, x in enumerate(qs):
                print(f"mfma graph rep={rep} input={i}", flush=True)
                r = ix.search(x, 10, 8, _z(10), mode=MODE_MFMA)
                np.testing.assert_array_equal(r[0], eager[i][0])
                np.testing.assert_array_equal(r[1], eager[i][1])
        assert ctx.info()["replays"] >= 5
    finally:
        ctx.set_graphs(False)
        ix.close()
