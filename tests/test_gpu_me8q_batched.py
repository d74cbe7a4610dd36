"""GPU tests of the batched ME8Q form (E8Q_MODE_BATCHED of e8q_m8_search).

The batched form computes the integer coarse sums of every (query, vector)
pair with an int8 MFMA contraction; they must equal the reference's
AccumulateCoarseHi sums (me8q_fast.cpp:800-818) exactly, which the checksum
test checks against a numpy restatement over the device's own LUT and
FastScan layout (both pinned to the reference's fixtures in
test_gpu_me8q.py).  With z = +inf nothing is pruned in either form, so the
batched result must be the plain form's exact estimator top-k; with the
reference's z the batched form applies the same bounds against a fixed
per-query threshold, and every distance it returns must be that vector's
exact estimator distance.
"""
import math

import numpy as np
import pytest

import capi
import capi_m8
import me8qx
from capi_m8 import BATCHED, PLAIN
from test_gpu_me8q import _oracle_case, _z

pytestmark = pytest.mark.gpu

INF = float("inf")


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


def _index(ctx, seed, n, dim, nlist, avg, spec="", sizes=None):
    data, cent, lab, q, oix = _oracle_case(seed, n, dim, nlist, avg, spec=spec, sizes=sizes)
    f = me8qx.parse(oix.image())
    ix = capi_m8.M8Index.from_file(ctx, f)
    return ix, f, q


def _run_cols(plan):
    cols, c0 = [], 0
    for ln, b in plan:
        if b > 0:
            cols.append((c0, c0 + ln // 8))
            c0 += ln // 8
    return cols


def _coarse_checksum(ix, f, q, nprobe_q):
    """numpy restatement: sum over the pairs of (pos + 1) * sum_r (r + 1) * S_r
    mod 2^32, S_r = sum over run r's plane columns of lut_lo[byte & 15] +
    lut_hi[byte >> 4]."""
    maxp = int(np.max(nprobe_q))
    _rq, lut64, _md, probe, _pd = ix.query(q, maxp)
    PC = ix.plane_cols
    lut64 = lut64.reshape(q.shape[0], PC, 64).astype(np.int64)
    lo, hi = lut64[:, :, 0:16], lut64[:, :, 32:48]
    runs = _run_cols(f["plan"])
    lb = np.asarray(f["list_begin"], np.int64)
    rec = PC * 32 + ix.nrun * 160
    planes = {}
    out = np.zeros(q.shape[0], np.int64)
    cidx = np.arange(PC)
    for qi in range(q.shape[0]):
        tot = 0
        for p in range(int(nprobe_q[qi])):
            c = int(probe[qi, p])
            cnt = int(lb[c + 1] - lb[c])
            if cnt == 0:
                continue
            if c not in planes:
                plane = ix.fast(c)[0].reshape(-1, rec)[:, :PC * 32].reshape(-1, PC, 32)
                planes[c] = plane.transpose(0, 2, 1).reshape(-1, PC)[:cnt].astype(np.int64)
            b = planes[c]
            vals = lo[qi][cidx, b & 15] + hi[qi][cidx, b >> 4]
            w = np.zeros(cnt, np.int64)
            for r, (c0, c1) in enumerate(runs):
                w += (r + 1) * vals[:, c0:c1].sum(axis=1)
            tot += int(np.sum((np.arange(cnt, dtype=np.int64) + 1) * w))
        out[qi] = tot % (1 << 32)
    return out.astype(np.uint32)


def _sorted_rows(ids, d):
    """Rows sorted by (distance, id), padding last."""
    out_i, out_d = ids.copy(), d.copy()
    for i in range(ids.shape[0]):
        o = np.lexsort((ids[i], d[i]))
        out_i[i], out_d[i] = ids[i][o], d[i][o]
    return out_i, out_d


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


CASES = [
    # seed, n, dim, nlist, avg, spec, sizes
    (31, 24000, 320, 48, 3.0, "64:8,128:4,128:2", None),
    (77, 0, 256, 10, 4.0, "64:6,128:4,64:2", np.array([3000, 5, 0, 700, 1, 64, 33, 2500, 31, 900])),
    (5, 6000, 128, 16, 4.0, "128:4", None),
    (9, 9000, 200, 20, 4.0, "64:9,64:5,64:3,64:1", None),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_coarse_sums_checksum(ctx, case):
    seed, n, dim, nlist, avg, spec, sizes = CASES[case]
    ix, f, q = _index(ctx, seed, n, dim, nlist, avg, spec, sizes)
    npq = np.random.default_rng(seed).integers(1, nlist + 1, q.shape[0]).astype(np.uint32)
    print(f"me8q batched checksum case {case}: plan {f['plan']}", flush=True)
    got = ix.batch_checksum(q, 10, nlist, _z(10, 0.0), nprobe_per_query=npq)
    want = _coarse_checksum(ix, f, q, npq)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("k,budget", [(10, 0), (100, 1), (7, 3)])
def test_unpruned_equals_plain(ctx, case, k, budget):
    seed, n, dim, nlist, avg, spec, sizes = CASES[case]
    ix, f, q = _index(ctx, seed, n, dim, nlist, avg, spec, sizes)
    nprobe = max(1, nlist // 3)
    print(f"me8q batched z=inf case {case} k={k} budget={budget} nprobe={nprobe}", flush=True)
    pi, pd, _, pst = ix.search(q, k, nprobe, INF)
    bi, bd, _, bst = ix.search(q, k, nprobe, INF, mode=BATCHED, budget=budget)
    assert bst.form == BATCHED and pst.form == PLAIN
    pi, pd = _sorted_rows(pi, pd)
    bi, bd = _sorted_rows(bi, bd)
    np.testing.assert_array_equal(bi, pi)
    # the cascade adds its run terms in fp64 in a per-group order; a float
    # result may move by one ulp between the forms
    assert int(np.max(_ulps(bd, pd))) <= 1
    assert bst.candidates == pst.candidates


@pytest.mark.parametrize("case", range(len(CASES)))
def test_pruned_distances_exact(ctx, case):
    seed, n, dim, nlist, avg, spec, sizes = CASES[case]
    ix, f, q = _index(ctx, seed, n, dim, nlist, avg, spec, sizes)
    k, nprobe = 10, max(2, nlist // 2)
    z = _z(k, 0.0)
    print(f"me8q batched pruned case {case}", flush=True)
    pi, pd, _, _ = ix.search(q, k, nprobe, z)
    bi, bd, _, bst = ix.search(q, k, nprobe, z, mode=BATCHED)
    assert bst.form == BATCHED
    # exact estimator distances of the probed vectors nearest each query
    ei, ed, _, _ = ix.search(q, 1024, nprobe, INF)
    found = 0
    for i in range(q.shape[0]):
        ex = {int(a): float(b) for a, b in zip(ei[i], ed[i]) if a != 0xFFFFFFFF}
        for a, b in zip(bi[i], bd[i]):
            if a == 0xFFFFFFFF:
                continue
            assert int(a) in ex, f"query {i}: id {a} is not among the probed vectors"
            assert _ulps(b, ex[int(a)]) <= 1, f"query {i}: id {a} distance {b} vs exact {ex[int(a)]}"
        # the batched k-th distance is never above the plain form's by more than rounding
        kb, kp = bd[i][bd[i] < INF], pd[i][pd[i] < INF]
        if kb.size == kp.size and kb.size > 0:
            assert kb[-1] <= kp[-1] + 1e-6 * max(1.0, abs(float(kp[-1])))
        found += len(set(map(int, bi[i])) & set(map(int, pi[i])))
    rec = found / float(q.shape[0] * k)
    print(f"  overlap with the plain form: {rec:.4f}", flush=True)
    assert rec >= 0.97


def test_uneven_nprobe_and_budget(ctx):
    sizes = np.array([3000, 5, 0, 700, 1, 64, 33, 2500, 31, 900])
    ix, f, q = _index(ctx, 77, 0, 256, 10, 4.0, "64:6,128:4,64:2", sizes)
    npq = np.random.default_rng(5).choice(np.array([1, 2, 3, 7, 10], np.uint32), size=q.shape[0])
    for budget in (1, 2, 5, 200):
        print(f"me8q batched uneven nprobe budget={budget}", flush=True)
        pi, pd, _, _ = ix.search(q, 20, 10, INF, nprobe_per_query=npq)
        bi, bd, _, bst = ix.search(q, 20, 10, INF, nprobe_per_query=npq, mode=BATCHED, budget=budget)
        pi, pd = _sorted_rows(pi, pd)
        bi, bd = _sorted_rows(bi, bd)
        np.testing.assert_array_equal(bi, pi)
        assert int(np.max(_ulps(bd, pd))) <= 1


def test_every_vector_a_candidate(ctx):
    # 40000 vectors in 4 lists, all probed, nothing pruned, and a threshold
    # pass shorter than k (tau = +inf): every vector is a candidate and a
    # survivor of its query
    ix, f, q = _index(ctx, 13, 40000, 128, 4, 4.0, "128:4")
    q = q[:8]
    print("me8q batched every vector", flush=True)
    pi, pd, _, _ = ix.search(q, 50, 4, INF)
    bi, bd, _, bst = ix.search(q, 50, 4, INF, mode=BATCHED, budget=1)
    assert bst.form == BATCHED and bst.overflowed == 0
    assert bst.refined >= q.shape[0] * (40000 - 32)
    pi, pd = _sorted_rows(pi, pd)
    bi, bd = _sorted_rows(bi, bd)
    np.testing.assert_array_equal(bi, pi)
    assert int(np.max(_ulps(bd, pd))) <= 1


@pytest.mark.parametrize("case", [0, 1, 3])
@pytest.mark.parametrize("cap", ["0", "100"])
def test_records_match_write_pass(ctx, monkeypatch, case, cap):
    # The count pass's records placed into the pool, against the write pass
    # (a second contraction) that runs when the records do not fit: the same
    # candidates with the same bounds and bins, so the same answer bit for bit.
    seed, n, dim, nlist, avg, spec, sizes = CASES[case]
    ix, f, q = _index(ctx, seed, n, dim, nlist, avg, spec, sizes)
    k, nprobe, z = 10, max(2, nlist // 2), _z(10, 0.0)
    print(f"me8q batched records vs write pass case {case} cap {cap}", flush=True)
    ri, rd, _, rst = ix.search(q, k, nprobe, z, mode=BATCHED)
    monkeypatch.setenv("E8Q_M8_RECORD_CAP", cap)
    wi, wd, _, wst = ix.search(q, k, nprobe, z, mode=BATCHED)
    monkeypatch.delenv("E8Q_M8_RECORD_CAP")
    assert rst.form == BATCHED and wst.form == BATCHED
    np.testing.assert_array_equal(ri, wi)
    np.testing.assert_array_equal(rd.view(np.uint32), wd.view(np.uint32))
    assert (rst.refined, rst.survivors, rst.pairs_scored) == (wst.refined, wst.survivors, wst.pairs_scored)


def test_unsupported_shapes_run_plain(ctx):
    ix, f, q = _index(ctx, 9, 3000, 384, 6, 4.0, "64:8,64:6,64:5,64:4,128:2")   # five coded runs
    print("me8q batched five runs", flush=True)
    pi, pd, _, _ = ix.search(q, 10, 3, _z(10, 0.0))
    bi, bd, _, bst = ix.search(q, 10, 3, _z(10, 0.0), mode=BATCHED)
    assert bst.form == PLAIN
    np.testing.assert_array_equal(bi, pi)
    np.testing.assert_array_equal(bd.view(np.uint32), pd.view(np.uint32))


def test_python_api_batched():
    import e8qlib

    rng = np.random.default_rng(3)
    n, dim, nlist = 8000, 192, 12
    cent = (rng.standard_normal((nlist, dim)) * 2.0).astype(np.float32)
    lab = rng.integers(0, nlist, n).astype(np.uint32)
    data = (cent[lab] + rng.standard_normal((n, dim))).astype(np.float32)
    q = (cent[rng.integers(0, nlist, 40)] + rng.standard_normal((40, dim))).astype(np.float32)
    idx = e8qlib.Me8qIndex(dim, 4.0, 0)
    idx.build(data, cent, lab, plan_spec="64:7,128:3")
    print("me8q batched python api", flush=True)
    pi, pd = idx.search(q, 10, 4)
    bi, bd, st = idx.search(q, 10, 4, return_stats=True, mode="batched")
    assert st["form"] == "batched"
    ov = np.mean([len(set(a) & set(b)) / 10.0 for a, b in zip(pi.tolist(), bi.tolist())])
    assert ov >= 0.97
    with pytest.raises(ValueError):
        idx.search(q, 10, 4, mode="mfma")
    assert math.isfinite(float(bd[0, 0]))


# ---------------------------------------------------------------------------
# The 10M x 1024 served configuration's plan (64:11, 128:6, 384:4, 448:2:
# a plane row of exactly 128 columns, the rate-11 level table staged in LDS)
# on Zipf-sized lists with a head list and empty lists, and a D = 1152 plan
# with 144 plane columns (the refine's word-loop row sums) and a rate-13 run
# (its level table read from global memory), against the oracle.

def _zipf_sizes(n, nlist, s=1.1, empty=(5, 17, 40)):
    w = np.arange(1, nlist + 1, dtype=np.float64) ** -s
    sizes = np.floor(w / w.sum() * n).astype(np.int64)
    for e in empty:
        if e < nlist:
            sizes[e] = 0
    return sizes


SHAPES = [
    # seed, n, dim, nlist, spec
    (101, 30000, 1024, 64, "64:11,128:6,384:4,448:2"),
    (102, 12000, 1152, 48, "64:13,192:5,448:3,448:2"),
]


@pytest.fixture(scope="module", params=range(len(SHAPES)), ids=["d1024_c5plan", "d1152_rate13"])
def shape_case(request, ctx):
    seed, n, dim, nlist, spec = SHAPES[request.param]
    data, cent, lab, q, oix = _oracle_case(seed, 0, dim, nlist, 4.0, spec=spec, sizes=_zipf_sizes(n, nlist))
    f = me8qx.parse(oix.image())
    ix = capi_m8.M8Index.from_file(ctx, f)
    assert [tuple(s) for s in f["plan"]] == [tuple(int(t) for t in s.split(":")) for s in spec.split(",")]
    return ix, oix, q, f, (data, cent, lab)


@pytest.mark.parametrize("k,nprobe", [(100, 8), (10, 24)])
def test_shape_unpruned_against_oracle(shape_case, k, nprobe):
    # z = +inf: nothing is pruned, the result is the exact estimator top-k
    # over the probed lists.  The plain form equals the oracle bit for bit;
    # the batched form has the same ids, and its distances are within one
    # ulp (its cascade adds the fp64 run terms in a group-of-one order).
    ix, oix, q, f, _ = shape_case
    print(f"me8q shape {f['dim']} unpruned k={k} nprobe={nprobe}", flush=True)
    oi, od, ost = oix.search(q, k, nprobe, z=INF)
    pi, pd, _, pst = ix.search(q, k, nprobe, INF)
    np.testing.assert_array_equal(pi, oi)
    np.testing.assert_array_equal(pd.view(np.uint32), od.view(np.uint32))
    assert pst.candidates == int(ost[:, 0].sum())
    bi, bd, _, bst = ix.search(q, k, nprobe, INF, mode=BATCHED)
    assert bst.form == BATCHED and bst.overflowed == 0
    assert bst.candidates == int(ost[:, 0].sum())
    oi, od = _sorted_rows(oi, od)
    bi, bd = _sorted_rows(bi, bd)
    np.testing.assert_array_equal(bi, oi)
    assert int(np.max(_ulps(bd, od))) <= 1


@pytest.mark.parametrize("k,nprobe", [(100, 8), (10, 24)])
def test_shape_pruned_against_oracle(shape_case, k, nprobe):
    # The reference's z: the plain form is the oracle bit for bit (ids,
    # distances, per-query counters); every distance the batched form
    # returns is that vector's exact estimator distance (the oracle's
    # unpruned search), and its k-th distance is never above the plain form's.
    ix, oix, q, f, _ = shape_case
    z = _z(k, 0.0)
    print(f"me8q shape {f['dim']} pruned k={k} nprobe={nprobe}", flush=True)
    oi, od, ost = oix.search(q, k, nprobe)
    pi, pd, pst, _ = ix.search(q, k, nprobe, z)
    np.testing.assert_array_equal(pi, oi)
    np.testing.assert_array_equal(pd.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(pst, ost)
    bi, bd, _, bst = ix.search(q, k, nprobe, z, mode=BATCHED)
    assert bst.form == BATCHED
    ei, ed, _ = oix.search(q, 1024, nprobe, z=INF)
    found = 0
    for i in range(q.shape[0]):
        ex = {int(a): float(b) for a, b in zip(ei[i], ed[i]) if a != 0xFFFFFFFF}
        for a, b in zip(bi[i], bd[i]):
            if a == 0xFFFFFFFF:
                continue
            assert int(a) in ex, f"query {i}: id {a} is not among the nearest probed vectors"
            assert _ulps(b, ex[int(a)]) <= 1, f"query {i}: id {a} distance {b} vs exact {ex[int(a)]}"
        kb, kp = bd[i][bd[i] < INF], pd[i][pd[i] < INF]
        if kb.size == kp.size and kb.size > 0:
            assert kb[-1] <= kp[-1] + 1e-6 * max(1.0, abs(float(kp[-1])))
        found += len(set(map(int, bi[i])) & set(map(int, oi[i])))
    rec = found / float(q.shape[0] * k)
    print(f"  overlap with the oracle's pruned search: {rec:.4f}", flush=True)
    assert rec >= 0.97


def test_shape_coarse_sums_checksum(shape_case):
    ix, oix, q, f, _ = shape_case
    npq = np.random.default_rng(3).integers(1, f["nlist"] + 1, q.shape[0]).astype(np.uint32)
    print(f"me8q shape {f['dim']} checksum", flush=True)
    got = ix.batch_checksum(q, 10, f["nlist"], _z(10, 0.0), nprobe_per_query=npq)
    np.testing.assert_array_equal(got, _coarse_checksum(ix, f, q, npq))
