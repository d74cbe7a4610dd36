"""Launch-independence of the batched ME8Q form (e8q_m8_search with
E8Q_MODE_BATCHED): graph replay (e8q_ctx_set_graphs), the persistent grids
and their autotuner (e8q_ctx_set_m8_grids), placement from the count pass's
records against the second contraction (E8Q_M8_RECORD_CAP), and the pool's
capacity (E8Q_M8_POOL_PER_QUERY, e8q_ctx_reserve).

Every per-query threshold of the batched form is fixed before the kernel
that applies it and each top-k is sorted by (distance, id), so none of
these may change a result: ids and distance bits must equal the eager
default launch's.  Queries whose candidates pass the pool's end are searched
by the plain form in the same device sequence; pool offsets grow with the
query index, so those are the last `overflowed` queries, and their results
must equal the plain form's bit for bit while the others keep the batched
results.
"""
import numpy as np
import pytest

import capi
from capi_m8 import BATCHED, PLAIN
import capi_m8
import me8qx
from test_gpu_me8q import _oracle_case, _z

pytestmark = pytest.mark.gpu

K, NPROBE = 20, 12
ZO = float("inf")   # the overflow tests: nothing pruned, every probed vector a candidate


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def case(ctx):
    _d, _c, _l, q, oix = _oracle_case(31, 24000, 320, 48, 3.0, spec="64:8,128:4,128:2")
    f = me8qx.parse(oix.image())
    ix = capi_m8.M8Index.from_file(ctx, f)
    rng = np.random.default_rng(3)
    qs = np.concatenate([q] + [(q + 0.05 * rng.standard_normal(q.shape)).astype(np.float32) for _ in range(3)])
    ctx.set_m8_grids(0, 0)
    ref = ix.search(qs, K, NPROBE, _z(K, 0.0), mode=BATCHED)
    assert ref[3].form == BATCHED and ref[3].overflowed == 0
    return ix, f, qs, ref


def _same(a, b, rows=slice(None)):
    np.testing.assert_array_equal(a[0][rows], b[0][rows])
    np.testing.assert_array_equal(a[1][rows].view(np.uint32), b[1][rows].view(np.uint32))


def test_grids_give_identical_results(ctx, case):
    ix, _f, qs, ref = case
    try:
        for g in [(1, 1), (7, 13), (256, 2048), (1024, 4096), (512, 64)]:
            ctx.set_m8_grids(*g)
            print(f"me8q batched grids {g}", flush=True)
            r = ix.search(qs, K, NPROBE, _z(K, 0.0), mode=BATCHED)
            info = ctx.m8_info()
            assert (info["mfma_grid"], info["refine_grid"]) == g
            _same(r, ref)
    finally:
        ctx.set_m8_grids(0, 0)


def test_grids_per_query_nprobe(ctx, case):
    ix, _f, qs, _ref = case
    npq = np.random.default_rng(4).choice(np.array([1, 3, 12, 48], np.uint32), size=qs.shape[0])
    try:
        ctx.set_m8_grids(0, 0)
        a = ix.search(qs, K, 48, _z(K, 0.0), nprobe_per_query=npq, mode=BATCHED)
        for g in [(3, 5), (600, 900)]:
            ctx.set_m8_grids(*g)
            print(f"me8q batched per-query nprobe grids {g}", flush=True)
            _same(ix.search(qs, K, 48, _z(K, 0.0), nprobe_per_query=npq, mode=BATCHED), a)
    finally:
        ctx.set_m8_grids(0, 0)


def test_autotuner_once_per_bucket(case):
    _ix, f, qs, ref = case
    c = capi.Ctx(0)   # a fresh context: nothing tuned yet
    try:
        ix = capi_m8.M8Index.from_file(c, f)
        print("me8q autotune first call", flush=True)
        r1 = ix.search(qs, K, NPROBE, _z(K, 0.0), mode=BATCHED)
        info = c.m8_info()
        assert info["tunings"] == 1 and info["mfma_grid"] > 0 and info["refine_grid"] > 0
        _same(r1, ref)
        print("me8q autotune same bucket", flush=True)
        r2 = ix.search(qs[:129], K, NPROBE, _z(K, 0.0), mode=BATCHED)   # 129 and 256 share a bucket
        assert c.m8_info()["tunings"] == 1
        _same(r2, ref, slice(0, 129))
        r3 = ix.search(qs[:10], K, NPROBE, _z(K, 0.0), mode=BATCHED)   # small: untuned default grids
        assert c.m8_info()["tunings"] == 1
        _same(r3, ref, slice(0, 10))
        r4 = ix.search(qs[:100], K, NPROBE, _z(K, 0.0), mode=BATCHED)   # another bucket
        assert c.m8_info()["tunings"] == 2
        _same(r4, ref, slice(0, 100))
        ix.close()
    finally:
        c.close()


def test_shared_tuning_table(case):
    """Two contexts sharing one table of tuned grids: a bucket the first
    tunes is not tuned again by the second, and both give the same results."""
    _ix, f, qs, ref = case
    c1, c2 = capi.Ctx(0), capi.Ctx(0)
    try:
        c2.share_tuning(c1)
        ix1 = capi_m8.M8Index.from_file(c1, f)
        ix2 = capi_m8.M8Index.from_file(c2, f)
        print("me8q shared tuning: first context tunes", flush=True)
        _same(ix1.search(qs, K, NPROBE, _z(K, 0.0), mode=BATCHED), ref)
        assert c1.m8_info()["tunings"] == 1
        print("me8q shared tuning: second context reuses", flush=True)
        _same(ix2.search(qs, K, NPROBE, _z(K, 0.0), mode=BATCHED), ref)
        assert c2.m8_info()["tunings"] == 0
        assert (c2.m8_info()["mfma_grid"], c2.m8_info()["refine_grid"]) == \
            (c1.m8_info()["mfma_grid"], c1.m8_info()["refine_grid"])
        r = ix2.search(qs[:100], K, NPROBE, _z(K, 0.0), mode=BATCHED)   # a bucket neither has seen
        assert c2.m8_info()["tunings"] == 1
        _same(r, ref, slice(0, 100))
        ix1.close()
        ix2.close()
    finally:
        c1.close()
        c2.close()


def test_graph_replay_matches_eager(ctx, case):
    ix, _f, qs, _ref = case
    rng = np.random.default_rng(5)
    ctx.set_graphs(False)
    inputs = [qs] + [(qs + 0.02 * rng.standard_normal(qs.shape)).astype(np.float32) for _ in range(3)]
    eager = [ix.search(x, K, NPROBE, _z(K, 0.0), mode=BATCHED) for x in inputs]
    other = ix.search(qs[:40], 5, 4, _z(5, 0.0), mode=BATCHED)
    ctx.set_graphs(True)
    try:
        before = ctx.info()
        for rep in range(3):
            for i, x in enumerate(inputs):
                print(f"me8q graph rep={rep} input={i}", flush=True)
                r = ix.search(x, K, NPROBE, _z(K, 0.0), mode=BATCHED)
                assert r[3].form == BATCHED
                _same(r, eager[i])
                if i % 2 == 1:   # eager work of another shape on the same context in between
                    _same(ix.search(qs[:40], 5, 4, _z(5, 0.0), mode=BATCHED), other)
        info = ctx.info()
        assert info["captures"] - before["captures"] >= 1
        assert info["replays"] - before["replays"] >= 8
    finally:
        ctx.set_graphs(False)


def test_graph_replay_per_query_nprobe(ctx, case):
    ix, _f, qs, _ref = case
    rng = np.random.default_rng(8)
    base = rng.choice(np.array([1, 2, 5, 16, 48], np.uint32), size=qs.shape[0])
    perms = [base] + [rng.permutation(base) for _ in range(3)]   # same pair count: one graph
    ctx.set_graphs(False)
    eager = [ix.search(qs, 10, 48, _z(10, 0.0), nprobe_per_query=n, mode=BATCHED) for n in perms]
    ctx.set_graphs(True)
    try:
        for rep in range(2):
            for i, n in enumerate(perms):
                print(f"me8q graph per-query nprobe rep={rep} perm={i}", flush=True)
                _same(ix.search(qs, 10, 48, _z(10, 0.0), nprobe_per_query=n, mode=BATCHED), eager[i])
    finally:
        ctx.set_graphs(False)


@pytest.mark.parametrize("cap", ["0", "1000"])
def test_write_pass_equals_records(ctx, case, monkeypatch, cap):
    ix, _f, qs, ref = case
    monkeypatch.setenv("E8Q_M8_RECORD_CAP", cap)
    print(f"me8q batched record cap {cap}: the write pass places", flush=True)
    _same(ix.search(qs, K, NPROBE, _z(K, 0.0), mode=BATCHED), ref)


def _check_overflow(r, ref, plain):
    nq = ref[0].shape[0]
    novf = int(r[3].overflowed)
    assert 0 < novf < nq
    _same(r, ref, slice(0, nq - novf))
    _same(r, plain, slice(nq - novf, nq))
    return novf


@pytest.fixture(scope="module")
def unpruned(ctx, case):
    """The case's batched and plain results with nothing pruned (z = +inf:
    every probed vector past the threshold pass is a candidate)."""
    ix, _f, qs, _ref = case
    return ix.search(qs, K, NPROBE, ZO, mode=BATCHED), ix.search(qs, K, NPROBE, ZO, mode=PLAIN)


@pytest.fixture(scope="module")
def half_pool(case):
    """Pool entries per query that hold about half of the case's candidates
    at z = +inf: a fresh context with a one-entry-per-query pool learns 1.25 x
    the candidates per query (plus one) from its first call (which then runs
    again at that size: nothing overflows in what it returns)."""
    _ix, f, qs, _ref = case
    mp = pytest.MonkeyPatch()
    mp.setenv("E8Q_M8_POOL_PER_QUERY", "1")
    c = capi.Ctx(0)
    try:
        ix = capi_m8.M8Index.from_file(c, f)
        r = ix.search(qs, K, NPROBE, ZO, mode=BATCHED)
        assert r[3].overflowed == 0
        learned = c.m8_info()["pool_per_query"]
        ix.close()
    finally:
        c.close()
        mp.undo()
    per_q = int((learned - 1) * 0.8 / 2)
    print(f"me8q batched: about {2 * per_q} candidates per query", flush=True)
    assert per_q >= 100
    return per_q


def test_pool_overflow_falls_back_to_plain(ctx, case, unpruned, half_pool, monkeypatch):
    _ix, f, qs, _ref = case
    ref, plain = unpruned
    monkeypatch.setenv("E8Q_M8_POOL_PER_QUERY", str(half_pool))
    c = capi.Ctx(0)   # an unreserved context that has learned nothing yet
    try:
        ix2 = capi_m8.M8Index.from_file(c, f)
        print(f"me8q batched pool of {half_pool} per query, unreserved", flush=True)
        # the call overflows, learns the need and runs again at that size:
        # the batched results for every query
        r = ix2.search(qs, K, NPROBE, ZO, mode=BATCHED)
        assert r[3].overflowed == 0
        _same(r, ref)
        assert c.m8_info()["pool_per_query"] > half_pool
        r2 = ix2.search(qs, K, NPROBE, ZO, mode=BATCHED)
        assert r2[3].overflowed == 0
        _same(r2, ref)
        ix2.close()
    finally:
        c.close()
    # a reserved context keeps its pool: the queries past it are answered by
    # the plain form inside the same sequence, eager and in a replayed graph
    # (it learns nothing)
    c4 = capi.Ctx(0)
    try:
        ix4 = capi_m8.M8Index.from_file(c4, f)
        c4.reserve(ix4.workspace_bytes(qs.shape[0], K, NPROBE, BATCHED))
        print("me8q batched overflow on a reserved context, eager", flush=True)
        novf = _check_overflow(ix4.search(qs, K, NPROBE, ZO, mode=BATCHED), ref, plain)
        print(f"  {novf} of {qs.shape[0]} queries overflowed", flush=True)
        assert c4.m8_info()["pool_per_query"] == 0
        ix4.close()
    finally:
        c4.close()
    # the same overflow inside a replayed graph, on a reserved context (which
    # learns nothing)
    c3 = capi.Ctx(0)
    try:
        ix3 = capi_m8.M8Index.from_file(c3, f)
        c3.reserve(ix3.workspace_bytes(qs.shape[0], K, NPROBE, BATCHED))
        c3.set_graphs(True)
        # the first call tunes the grids, the second runs eagerly, then
        # captures and replays
        for rep in range(4):
            print(f"me8q graph with overflow rep={rep}", flush=True)
            _check_overflow(ix3.search(qs, K, NPROBE, ZO, mode=BATCHED), ref, plain)
        assert c3.info()["replays"] >= 2
        assert c3.m8_info()["pool_per_query"] == 0
        ix3.close()
    finally:
        c3.close()


def test_reserved_workspace_is_never_grown(case, unpruned, half_pool, monkeypatch):
    _ix, f, qs, _ref = case
    ref, plain = unpruned
    nq = qs.shape[0]
    c = capi.Ctx(0)
    try:
        ix = capi_m8.M8Index.from_file(c, f)
        need = ix.workspace_bytes(nq, K, NPROBE, BATCHED)
        assert need > ix.workspace_bytes(nq, K, NPROBE, PLAIN)
        monkeypatch.setenv("E8Q_M8_POOL_PER_QUERY", str(half_pool))
        small = ix.workspace_bytes(nq, K, NPROBE, BATCHED)
        monkeypatch.delenv("E8Q_M8_POOL_PER_QUERY")
        assert small < need
        c.reserve(need)
        ws = c.workspace()
        print(f"me8q batched in a reserved workspace of {need} bytes", flush=True)
        r = ix.search(qs, K, NPROBE, ZO, mode=BATCHED)
        assert r[3].overflowed == 0 and c.workspace() == ws
        _same(r, ref)
        r = ix.search(qs[:50], K, 4, ZO, mode=BATCHED)   # smaller calls fit too
        assert c.workspace() == ws
        ix.close()
    finally:
        c.close()
    # a reservation that holds about half the candidates: the pool shrinks to
    # fit it, the queries past it go to the plain form, the workspace stays
    c2 = capi.Ctx(0)
    try:
        ix2 = capi_m8.M8Index.from_file(c2, f)
        c2.reserve(small)
        ws2 = c2.workspace()
        print(f"me8q batched in a reserved workspace of {ws2} bytes", flush=True)
        r = ix2.search(qs, K, NPROBE, ZO, mode=BATCHED)
        assert c2.workspace() == ws2
        _check_overflow(r, ref, plain)
        assert c2.m8_info()["pool_per_query"] == 0
        ix2.close()
    finally:
        c2.close()


def test_reserved_small_call_takes_the_whole_reservation(case, unpruned, half_pool, monkeypatch):
    """A reservation sized for every query at a per-query default that holds
    half of each one's candidates: a call of a few queries gets all of it for
    its pool, so nothing overflows (a serving worker's small batch with a
    heavy query is answered by the batched form, not the plain one)."""
    _ix, f, qs, _ref = case
    ref, _plain = unpruned
    monkeypatch.setenv("E8Q_M8_POOL_PER_QUERY", str(half_pool))
    c = capi.Ctx(0)
    try:
        ix = capi_m8.M8Index.from_file(c, f)
        c.reserve(ix.workspace_bytes(qs.shape[0], K, NPROBE, BATCHED))
        ws = c.workspace()
        for n in (1, 4, 9):
            print(f"me8q batched: {n} queries in a reservation for {qs.shape[0]}", flush=True)
            r = ix.search(qs[:n], K, NPROBE, ZO, mode=BATCHED)
            assert r[3].overflowed == 0 and c.workspace() == ws
            _same(r, ref, slice(0, n))
        assert c.m8_info()["pool_per_query"] == 0
        ix.close()
    finally:
        c.close()


def test_grid_validation(ctx):
    with pytest.raises(RuntimeError):
        ctx.set_m8_grids(70000, 0)
