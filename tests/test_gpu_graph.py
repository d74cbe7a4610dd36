"""Graph replay and the batched-scan shape (e8q_ctx_set_graphs,
e8q_ctx_set_batched_shape and the context's autotuner).

Results must not depend on how a search is launched: every batched scan
shape (pairs and batches per work item, persistent grid) gives the same ids
and distances bit for bit, the tuned shape included, and a replayed graph
gives what the eager launches give for the same inputs.  Graph replays run
several times with other eager searches on the same context in between, with
fresh query values and per-query nprobe, and through the overflow fallback
that runs after the graph.
"""
import numpy as np
import pytest

import capi
from test_gpu_batched import _mixture_index, _z

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def mix(ctx):
    return _mixture_index(ctx, 30000, 128, 48, 4, 21, 300)


def _same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("disable_prune", [False, True])
def test_shapes_give_identical_results(ctx, mix, disable_prune):
    ix, q = mix
    ctx.set_batched_shape(0, 0, 0)
    ref = None
    # (qtile, btile, grid): default tiles, tiny and large items, one workgroup
    for shape in [(8, 16, 0), (4, 2, 0), (32, 32, 0), (16, 8, 1), (7, 6, 300)]:
        ctx.set_batched_shape(*shape)
        print(f"batched shape {shape} disable_prune={disable_prune}", flush=True)
        r = ix.search(q, 20, 12, _z(20), disable_prune=disable_prune, mode=capi.MODE_BATCHED)
        assert ctx.info()["qtile"] == shape[0] and ctx.info()["btile"] == shape[1]
        if ref is None:
            ref = r
        else:
            _same(r, ref)
    ctx.set_batched_shape(0, 0, 0)


def test_shape_validation(ctx):
    with pytest.raises(RuntimeError):
        ctx.set_batched_shape(8, 3, 0)   # odd btile
    with pytest.raises(RuntimeError):
        ctx.set_batched_shape(8, 0, 0)   # tiles set together


def test_autotuner_picks_once_per_bucket(ctx, mix):
    ix0, q = mix
    ctx.set_batched_shape(8, 16, 0)
    fixed = ix0.search(q, 10, 16, _z(10), mode=capi.MODE_BATCHED)
    ctx.set_batched_shape(0, 0, 0)
    c = capi.Ctx(0)   # a fresh context: nothing tuned yet
    try:
        ix, q2 = _mixture_index(c, 30000, 128, 48, 4, 21, 300)   # the same seeded index
        np.testing.assert_array_equal(q, q2)
        print("autotune first call", flush=True)
        r1 = ix.search(q, 10, 16, _z(10), mode=capi.MODE_BATCHED)
        info = c.info()
        assert info["tunings"] == 1
        assert info["qtile"] > 0 and info["btile"] >= 2 and info["grid"] > 0
        _same(r1, fixed)
        print("autotune same bucket", flush=True)
        r2 = ix.search(q[:257], 10, 16, _z(10), mode=capi.MODE_BATCHED)   # 257 and 300 share the 512 bucket
        r3 = ix.search(q, 10, 16, _z(10), mode=capi.MODE_BATCHED)
        assert c.info()["tunings"] == 1
        _same(r3, fixed)
        _same(r2, (fixed[0][:257], fixed[1][:257]))
        # a small call runs untuned at the default shape
        r4 = ix.search(q[:10], 10, 16, _z(10), mode=capi.MODE_BATCHED)
        assert c.info()["tunings"] == 1
        _same(r4, (fixed[0][:10], fixed[1][:10]))
        # another bucket tunes again
        r5 = ix.search(q[:100], 10, 16, _z(10), mode=capi.MODE_BATCHED)
        assert c.info()["tunings"] == 2
        _same(r5, (fixed[0][:100], fixed[1][:100]))
        ix.close()
    finally:
        c.close()


@pytest.mark.parametrize("mode", [capi.MODE_PLAIN, capi.MODE_BATCHED])
def test_graph_replay_matches_eager(ctx, mix, mode):
    ix, q = mix
    rng = np.random.default_rng(5)
    ctx.set_graphs(False)
    qs = [q] + [(q + 0.05 * rng.standard_normal(q.shape)).astype(np.float32) for _ in range(4)]
    eager = [ix.search(x, 10, 16, _z(10), mode=mode) for x in qs]
    other = ix.search(q[:50], 5, 4, _z(5), mode=mode)
    ctx.set_graphs(True)
    try:
        before = ctx.info()
        for rep in range(3):
            for i, x in enumerate(qs):
                print(f"graph mode={mode} rep={rep} input={i}", flush=True)
                _same(ix.search(x, 10, 16, _z(10), mode=mode), eager[i])
                # eager work of another shape on the same context in between
                if i % 2 == 1:
                    _same(ix.search(q[:50], 5, 4, _z(5), mode=mode), other)
        info = ctx.info()
        assert info["captures"] - before["captures"] >= 1
        assert info["replays"] - before["replays"] >= 12
    finally:
        ctx.set_graphs(False)


def test_graph_replay_per_query_nprobe(ctx, mix):
    ix, q = mix
    rng = np.random.default_rng(8)
    base = rng.choice(np.array([1, 2, 5, 16, 48], np.uint32), size=q.shape[0])
    # permutations keep the pair count (one graph), the values move between queries
    perms = [base] + [rng.permutation(base) for _ in range(3)]
    ctx.set_graphs(False)
    eager = [ix.search(q, 10, 48, _z(10), nprobe_per_query=npq, mode=capi.MODE_BATCHED) for npq in perms]
    ctx.set_graphs(True)
    try:
        for rep in range(2):
            for i, npq in enumerate(perms):
                print(f"graph per-query nprobe rep={rep} perm={i}", flush=True)
                _same(ix.search(q, 10, 48, _z(10), nprobe_per_query=npq, mode=capi.MODE_BATCHED), eager[i])
    finally:
        ctx.set_graphs(False)


def test_graph_replay_with_overflow_fallback(ctx):
    # Pruning off on a skewed index: more candidates than a query keeps, so
    # the plain search runs again after the graph for those queries.
    sizes = np.array([6000, 40, 900, 3, 1200, 0, 700, 350], np.int64)
    ix, q = _mixture_index(ctx, 0, 64, 8, 4, 8, 60, sizes=sizes)
    ctx.set_graphs(False)
    eager = ix.search(q, 50, 8, _z(50), disable_prune=True, mode=capi.MODE_BATCHED)
    plain = ix.search(q, 50, 8, _z(50), disable_prune=True)
    _same(eager, plain)
    ctx.set_graphs(True)
    try:
        for rep in range(4):
            print(f"graph overflow rep={rep}", flush=True)
            _same(ix.search(q, 50, 8, _z(50), disable_prune=True, mode=capi.MODE_BATCHED), eager)
            _same(ix.search(q[:7], 50, 8, _z(50), disable_prune=True), (plain[0][:7], plain[1][:7]))
    finally:
        ctx.set_graphs(False)
        ix.close()
