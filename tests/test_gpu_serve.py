"""The serving loop on the device: requests of mixed sizes, k and nprobe
submitted concurrently from several threads are batched together, and every
request's ids and distances equal the same request searched alone (bit for
bit) -- for an E8Q and an ME8Q index.  Admission by memory rejects a request
that cannot fit a worker's workspace."""
import os
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mixed_requests(base_q, n_req, seed, ks, nprobes):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_req):
        nq = int(rng.integers(1, 24))
        rows = rng.integers(0, base_q.shape[0], nq)
        q = (base_q[rows] + 0.05 * rng.standard_normal((nq, base_q.shape[1]))).astype(np.float32)
        out.append((q, int(rng.choice(ks)), int(rng.choice(nprobes))))
    return out


def _serve_and_compare(idx, reqs, workers, mode="plain"):
    import e8qlib
    srv = e8qlib.Server(idx, workers=workers, max_batch=256, mode=mode)
    tickets = [None] * len(reqs)

    def client(part):
        for i in range(part, len(reqs), 3):
            q, k, npb = reqs[i]
            tickets[i] = srv.submit(q, k, npb)

    th = [threading.Thread(target=client, args=(p,)) for p in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    results = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    for (q, k, npb), (ids, d, info) in zip(reqs, results):
        ai, ad = idx.search(q, k, npb) if mode == "plain" else idx.search(q, k, npb, mode=mode)
        np.testing.assert_array_equal(ids, ai)
        np.testing.assert_array_equal(d.view(np.uint32), ad.view(np.uint32))
    assert st["batches"] < len(reqs), st
    assert st["max_workspace_used"] <= st["workspace_bytes"]
    return st


def test_e8q_served_equals_alone():
    import e8qlib
    idx = e8qlib.E8qIndex.load(os.path.join(G, "d128b4.e8ax"))
    base_q = np.load(os.path.join(G, "d128b4_search.npz"))["queries"]
    print("serve e8q", flush=True)
    _serve_and_compare(idx, _mixed_requests(base_q, 60, 1, (10, 100), (1, 3, 8, 17)), workers=3)


def test_me8q_served_equals_alone():
    import e8qlib
    idx = e8qlib.Me8qIndex.load(os.path.join(G, "m400a4.me8q"))
    base_q = np.load(os.path.join(G, "m400a4_me8q.npz"))["queries"]
    print("serve me8q", flush=True)
    _serve_and_compare(idx, _mixed_requests(base_q, 60, 2, (10, 100), (1, 2, 5, 9)), workers=2)


def test_me8q_batched_served_equals_alone():
    # The batched ME8Q form (the served 10M x 1024 configuration's form) at
    # that configuration's plan on Zipf-sized lists: each request's ids and
    # distance bits are the same alone and in the mix (every threshold and
    # candidate split of the form is per query).
    import e8qlib
    from test_gpu_me8q import _oracle_case
    from test_gpu_me8q_batched import _zipf_sizes
    data, cent, lab, q, _oix = _oracle_case(101, 0, 1024, 64, 4.0, spec="64:11,128:6,384:4,448:2",
                                            sizes=_zipf_sizes(30000, 64))
    idx = e8qlib.Me8qIndex(1024, 4.0, 0)
    idx.build(data, cent, lab, plan_spec="64:11,128:6,384:4,448:2")
    print("serve me8q batched", flush=True)
    reqs = _mixed_requests(q, 48, 3, (10, 100), (1, 3, 8, 17))
    st = _serve_and_compare(idx, reqs, workers=2, mode="batched")
    print(f"  {st}", flush=True)
    # the batched form ran: its stats name it
    _i, _d, s = idx.search(reqs[0][0], reqs[0][1], reqs[0][2], return_stats=True, mode="batched")
    assert s["form"] == "batched"


def test_admission_rejects_oversized_request():
    import e8qlib
    idx = e8qlib.E8qIndex.load(os.path.join(G, "d128b4.e8ax"))
    srv = e8qlib.Server(idx, workers=1, max_batch=64)
    with pytest.raises(ValueError):
        srv.submit(np.zeros((65, 128), np.float32), 10, 4)
    ids, _, _ = srv.wait(srv.submit(np.zeros((64, 128), np.float32), 10, 4))
    assert ids.shape == (64, 10)
    srv.close()
