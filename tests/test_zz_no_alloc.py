"""The allocation-free plain search (e8q_search_static, include/e8q_gpu.h)
against e8q_search, through the C ABI.

Both run the same kernels; the static entry point only takes its buffers from
the fixed per-GPU scratch array instead of the context workspace.  So ids,
distance bits and every counter must be equal, on consecutive calls whose
second call is smaller than the first (the array keeps the first call's
contents), with host and with device queries and outputs, and the context
that only ran the static entry point must hold no workspace afterwards.
"""
import numpy as np
import pytest

import capi
import oracle_py as orc
from tables import empty_arrays, layout, tables

pytestmark = pytest.mark.gpu

LIMIT = 32 << 20   # E8Q_STATIC_SCRATCH_BYTES


def _index(ctx, seed=21, n=30000, dim=128, nlist=96, bits=4):
    rng = np.random.default_rng(seed)
    cent = (rng.standard_normal((nlist, dim)) * 2.0).astype(np.float32)
    lab = rng.integers(0, nlist, n).astype(np.uint32)
    data = (cent[lab] + rng.standard_normal((n, dim)).astype(np.float32)).astype(np.float32)
    q = (cent[rng.integers(0, nlist, 3000)] + rng.standard_normal((3000, dim)).astype(np.float32)).astype(np.float32)
    lay = layout(lab, nlist, bits, dim)
    arr = empty_arrays(n, dim, nlist, bits, lay, orc.rotator_signs(dim, seed))
    ix = capi.DeviceIndex(ctx, arr, tables(bits, lay["padded"]))
    ix.encode(data, cent, lab, lay["slot"])
    return ix, q


def _z(k):
    return orc.lib().orc_ortho_quantile(1.0, max(4 * k, 32))


def _same(a, b):
    ai, ad, ast = a
    bi, bd, bst = b
    np.testing.assert_array_equal(ai, bi)
    np.testing.assert_array_equal(ad.view(np.uint32), bd.view(np.uint32))
    for f in ("candidates", "survivors", "batches", "scanned"):
        assert ast[f] == bst[f], f


def test_static_equals_workspace_search():
    ctx = capi.Ctx(0)
    try:
        ix, q = _index(ctx)
        assert capi.lib().e8q_ctx_workspace_bytes(ctx.h) == 0
        # first call: 3000 queries, k=100, nprobe=24; second: 37 queries,
        # k=10, per-query nprobe (smaller in every buffer)
        npq = np.random.default_rng(4).integers(1, 40, 37).astype(np.uint32)
        calls = [dict(queries=q, k=100, nprobe=24, z=_z(100)),
                 dict(queries=q[:37], k=10, nprobe=0, z=_z(10), nprobe_per_query=npq),
                 dict(queries=q[100:1100], k=50, nprobe=96, z=_z(50), disable_prune=True, flush_every=7)]
        got = []
        for i, c in enumerate(calls):
            assert ix.workspace_bytes(c["queries"].shape[0], c["k"], max(c["nprobe"], 40)) <= LIMIT
            print(f"static search call {i}: nq={c['queries'].shape[0]} k={c['k']}", flush=True)
            got.append(ix.search_full(static=True, **c))
        # the static entry point grew no workspace
        assert capi.lib().e8q_ctx_workspace_bytes(ctx.h) == 0
        for i, c in enumerate(calls):
            print(f"workspace search call {i}", flush=True)
            _same(got[i], ix.search_full(**c))
    finally:
        ctx.close()


def test_static_device_pointers():
    import torch
    ctx = capi.Ctx(0)
    try:
        ix, q = _index(ctx, seed=5, n=8000, nlist=40)
        for nq, k, nprobe in ((1500, 64, 16), (9, 5, 3)):
            print(f"static search, device buffers: nq={nq} k={k}", flush=True)
            qd = torch.from_numpy(q[:nq]).to("cuda:0")
            oid = torch.empty((nq, k), dtype=torch.int32, device="cuda:0")
            od = torch.empty((nq, k), dtype=torch.float32, device="cuda:0")
            torch.cuda.synchronize()
            _, _, st = ix.search_full(nq, k, nprobe, _z(k), static=True,
                                      device_ptrs=(qd.data_ptr(), oid.data_ptr(), od.data_ptr()))
            wi, wd, wst = ix.search_full(q[:nq], k, nprobe, _z(k))
            _same((oid.cpu().numpy().view(np.uint32), od.cpu().numpy(), st), (wi, wd, wst))
    finally:
        ctx.close()


def test_static_rejects_what_does_not_fit():
    ctx = capi.Ctx(0)
    try:
        ix, q = _index(ctx, seed=8, n=4000, nlist=2048)
        # 2048 route distances + probe lists per query: 6000 queries need
        # more than the array holds
        big = np.repeat(q, 2, axis=0)
        assert ix.workspace_bytes(big.shape[0], 10, 64) > LIMIT
        with pytest.raises(RuntimeError, match="allocation-free entry point holds"):
            ix.search_full(big, 10, 64, _z(10), static=True)
        with pytest.raises(RuntimeError, match="plain form only"):
            s = capi.SearchArgs()
            qq = np.ascontiguousarray(q[:4])
            ids = np.empty((4, 10), np.uint32)
            dd = np.empty((4, 10), np.float32)
            s.queries, s.nq, s.k, s.nprobe, s.z = capi.p(qq), 4, 10, 4, _z(10)
            s.mode, s.out_ids, s.out_dists, s.flush_every = capi.MODE_BATCHED, capi.p(ids), capi.p(dd), 32
            capi.check(capi.lib().e8q_search_static(ctx.h, ix.h, capi.ctypes.byref(s), None))
        # a failed call leaves the entry point usable
        _same(ix.search_full(q[:20], 10, 8, _z(10), static=True), ix.search_full(q[:20], 10, 8, _z(10)))
    finally:
        ctx.close()
