"""The serving loop's scheduler (e8q-artifact_amd/csrc/host/serve.cpp) on
host-only workers: results go to the right request, batches mix requests of
one k only, never exceed the batch limit or the workspace (admission by
memory), a request that can never fit is rejected, the oldest waiting
request heads every batch, and close() drains the queue.  No GPU needed."""
import threading

import numpy as np
import pytest

import e8qlib

DIM = 8


def _expect(q, k, nprobe):
    bits = q[:, 0].astype(np.float32).view(np.uint32).astype(np.uint64)
    ids = (bits[:, None] + np.arange(k, dtype=np.uint64)[None, :] + nprobe).astype(np.uint32)
    d = q[:, :1].astype(np.float32) + np.arange(k, dtype=np.float32)[None, :]
    return ids, d


def _req(rng, nq):
    return rng.standard_normal((nq, DIM)).astype(np.float32)


def test_results_routed_and_batches_formed():
    srv = e8qlib.Server._sim(DIM, 64, 2000.0, workers=2, max_batch=64)
    rng = np.random.default_rng(0)
    reqs = []
    lock = threading.Lock()

    def client(seed):
        r = np.random.default_rng(seed)
        for _ in range(40):
            q = _req(r, int(r.integers(1, 9)))
            k = int(r.choice([3, 10]))
            npb = int(r.integers(1, 30))
            t = srv.submit(q, k, npb)
            with lock:
                reqs.append((t, q, k, npb))

    th = [threading.Thread(target=client, args=(s,)) for s in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    multi = 0
    for t, q, k, npb in reqs:
        ids, d, info = srv.wait(t)
        ei, ed = _expect(q, k, npb)
        np.testing.assert_array_equal(ids, ei)
        np.testing.assert_array_equal(d, ed)
        assert info["batch_queries"] >= q.shape[0]
        multi += info["batch_requests"] > 1
    st = srv.stats()
    srv.close()
    assert st["requests"] == len(reqs) and st["batches"] < len(reqs)
    assert st["max_batch_queries"] <= 64
    assert multi > 0   # continuous batching packed requests together
    del rng


def test_admission_by_memory():
    # workspace per query = 100 + 8k + 4*nprobe bytes; budget 10 kB
    srv = e8qlib.Server._sim(DIM, 100, 500.0, workers=1, max_batch=1000, workspace_bytes=10_000)
    rng = np.random.default_rng(1)
    with pytest.raises(ValueError, match="memory budget"):
        srv.submit(_req(rng, 50), 10, 8)          # 50 * (100 + 80 + 32) > 10000
    ok = [srv.submit(_req(rng, 10), 10, 8) for _ in range(30)]   # each fits alone
    for t in ok:
        srv.wait(t)
    st = srv.stats()
    srv.close()
    assert st["rejected"] == 1
    assert st["max_workspace_used"] <= 10_000
    assert st["batches"] >= 30 * 10 * 212 // 10_000


def test_oldest_request_heads_each_batch():
    srv = e8qlib.Server._sim(DIM, 8, 30000.0, workers=1, max_batch=1000)
    rng = np.random.default_rng(2)
    first = srv.submit(_req(rng, 1), 5, 1)         # occupies the worker
    a = srv.submit(_req(rng, 2), 10, 1)
    b = srv.submit(_req(rng, 3), 20, 1)
    c = srv.submit(_req(rng, 4), 10, 1)
    srv.wait(first)
    _, _, ia = srv.wait(a)
    _, _, ic = srv.wait(c)
    _, _, ib = srv.wait(b)
    srv.close()
    # a heads the second batch and c (same k) joins it; b runs after them
    assert ia["batch_requests"] == 2 and ic["batch_requests"] == 2 and ia["batch_queries"] == 6
    assert ib["batch_requests"] == 1
    assert ib["queue_us"] > ia["queue_us"]


def test_close_drains_and_rejects_after():
    srv = e8qlib.Server._sim(DIM, 8, 1000.0, workers=2, max_batch=16)
    rng = np.random.default_rng(3)
    ts = [srv.submit(_req(rng, 3), 4, 2) for _ in range(20)]
    srv.close()
    for t in ts:
        ids, _, _ = srv.wait(t)
        assert ids.shape == (3, 4)
    with pytest.raises(RuntimeError):
        srv.submit(_req(rng, 1), 4, 2)


def test_bad_arguments():
    srv = e8qlib.Server._sim(DIM, 8, 0.0, workers=1)
    with pytest.raises(ValueError):
        srv.submit(np.zeros((1, DIM), np.float32), 0, 1)
    with pytest.raises(ValueError):
        srv.submit(np.zeros((1, DIM), np.float32), 3, 0)
    with pytest.raises(ValueError):
        srv.wait(12345)
    srv.close()
