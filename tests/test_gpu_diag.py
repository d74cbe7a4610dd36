"""The diagnostics on the device, through the C ABI (e8q_estimate,
e8q_probe_bound_w, e8q_exhaustive_topk, e8q_m8_estimate,
e8q_m8_exhaustive_topk, e8q_m8_search_ivf): against fixtures written by the
reference library itself (tests/golden/diag.npz, the reference's
E8qIndex::Estimate / ProbeBoundW / ExhaustiveTopK and Me8qIndex::Estimate /
ExhaustiveTopK / SearchIVF on every fixture index), and at larger sizes
against the oracle, which the same fixtures pin (tests/test_oracle_diag.py).

Everything is compared exactly: float and double bits (and the NaN pattern
of ProbeBoundW), id lists in order -- the ME8Q lists in PushCandidate's
insertion order, ties included (m64drop at k = 400) -- and SearchIVF's
counters.
"""
import os

import numpy as np
import pytest

import capi
import capi_m8
import cases
import e8ax
import me8qx
import oracle_py as orc
from tables import empty_arrays, layout, tables

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(G, "diag.npz"))


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _same_f(got, want):
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_array_equal(_bits(got[ok]), _bits(want[ok]))


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_e8q_diagnostics_against_reference(ctx, gold, name):
    data, _c, _l, queries = cases.mixture(name)
    a = e8ax.read(os.path.join(G, f"{name}.e8ax"))
    ix = capi.DeviceIndex(ctx, a, tables(a["bits"], a["padded"]))
    ix.upload()
    pq, pid = cases.diag_pairs(name)
    print(f"e8q diagnostics {name}: estimate, probe_bound_w", flush=True)
    _same_f(ix.estimate(queries, pq, pid), gold[f"{name}_est"])
    _same_f(ix.probe_bound_w(queries, data[pid], pq, pid), gold[f"{name}_pbw"])
    for k in cases.diag_k(name):
        print(f"e8q diagnostics {name}: exhaustive k={k}", flush=True)
        ids, d = ix.exhaustive_topk(queries, k)
        np.testing.assert_array_equal(ids, gold[f"{name}_k{k}_exids"])
        np.testing.assert_array_equal(_bits(d), _bits(gold[f"{name}_k{k}_exd"]))


@pytest.mark.parametrize("name", list(cases.ME8Q_CASES))
def test_me8q_diagnostics_against_reference(ctx, gold, name):
    _d, _c, _l, queries, _lam = cases.me8q_mixture(name)
    ix = capi_m8.M8Index.from_file(ctx, me8qx.parse(np.fromfile(os.path.join(G, f"{name}.me8q"), np.uint8)))
    pq, pid = cases.diag_pairs(name)
    print(f"me8q diagnostics {name}: estimate", flush=True)
    np.testing.assert_array_equal(_bits(ix.estimate(queries, pq, pid)), _bits(gold[f"{name}_est"]))
    for k in sorted({k for k, _ in cases.diag_ivf_grid(name)}):
        print(f"me8q diagnostics {name}: exhaustive k={k}", flush=True)
        ids, d = ix.exhaustive_topk(queries, k)
        np.testing.assert_array_equal(ids, gold[f"{name}_k{k}_exids"])
        np.testing.assert_array_equal(_bits(d), _bits(gold[f"{name}_k{k}_exd"]))
    for (k, nprobe) in cases.diag_ivf_grid(name):
        print(f"me8q diagnostics {name}: search_ivf k={k} nprobe={nprobe}", flush=True)
        ids, d, st = ix.search_ivf(queries, k, nprobe)
        tag = f"{name}_k{k}_p{nprobe}"
        np.testing.assert_array_equal(ids, gold[f"{tag}_ivfids"])
        np.testing.assert_array_equal(_bits(d), _bits(gold[f"{tag}_ivfd"]))
        np.testing.assert_array_equal(st, gold[f"{tag}_ivfstats"])
    ix.close()


def test_e8q_errors_and_out_of_range(ctx):
    data, _c, _l, queries = cases.mixture("d128b4")
    a = e8ax.read(os.path.join(G, "d128b4.e8ax"))
    ix = capi.DeviceIndex(ctx, a, tables(a["bits"], a["padded"]))
    ix.upload()
    with pytest.raises(RuntimeError, match="id out of range"):
        ix.estimate(queries, [0], [3000])
    with pytest.raises(RuntimeError, match="query index out of range"):
        ix.estimate(queries, [queries.shape[0]], [1])
    # the batch form's out-of-range pairs give NaN, the rest their values
    pbw = ix.probe_bound_w(queries, data[[7, 7, 8]], [0, queries.shape[0], 1], [7, 7, 3000 + 1])
    assert not np.isnan(pbw[0]) and np.isnan(pbw[1]) and np.isnan(pbw[2])
    with pytest.raises(RuntimeError, match="k must be"):
        ix.exhaustive_topk(queries, 0)
    ids, d = ix.exhaustive_topk(queries[:2], 1024)   # k above n is not reached here: n = 3000
    assert ids.shape == (2, 1024) and np.all(np.diff(d, axis=1) >= 0)


def test_me8q_errors_and_empty_calls(ctx):
    _d, _c, _l, queries, _lam = cases.me8q_mixture("m64drop")
    ix = capi_m8.M8Index.from_file(ctx, me8qx.parse(np.fromfile(os.path.join(G, "m64drop.me8q"), np.uint8)))
    with pytest.raises(RuntimeError, match="unknown id"):
        ix.estimate(queries, [0], [300])
    ids, d = ix.exhaustive_topk(queries, 1000)   # k above n: padded
    assert np.all(ids[:, 300:] == 0xFFFFFFFF) and np.all(np.isinf(d[:, 300:]))
    assert np.all(ids[:, :300] != 0xFFFFFFFF)
    ix.close()


def test_e8q_diagnostics_against_oracle_at_size(ctx):
    rng = np.random.default_rng(21)
    n, dim, nlist, bits = 30000, 256, 48, 4
    cent = (rng.standard_normal((nlist, dim)) * 2.0).astype(np.float32)
    lab = rng.integers(0, nlist, n).astype(np.uint32)
    data = (cent[lab] + rng.standard_normal((n, dim)).astype(np.float32)).astype(np.float32)
    data[100:140] = data[99]   # equal vectors: equal distances, the tie order decided by id
    q = (cent[rng.integers(0, nlist, 24)] + rng.standard_normal((24, dim)).astype(np.float32)).astype(np.float32)
    q[0] = data[99]
    oi = orc.OracleIndex.build(data, cent, lab, bits, 42)
    lay = layout(lab, nlist, bits, dim)
    arr = empty_arrays(n, dim, nlist, bits, lay, orc.rotator_signs(dim, 42))
    ix = capi.DeviceIndex(ctx, arr, tables(bits, lay["padded"]))
    print("e8q diagnostics at size: build", flush=True)
    ix.encode(data, cent, lab, lay["slot"])
    pq = rng.integers(0, q.shape[0], 600).astype(np.uint32)
    pid = rng.integers(0, n, 600).astype(np.uint32)
    print("e8q diagnostics at size: estimate, probe_bound_w", flush=True)
    _same_f(ix.estimate(q, pq, pid), oi.estimate(q, pq, pid))
    _same_f(ix.probe_bound_w(q, data[pid], pq, pid), oi.probe_bound_w(q, data[pid], pq, pid))
    print("e8q diagnostics at size: exhaustive k=100", flush=True)
    ids, d = ix.exhaustive_topk(q, 100)
    oids, od = oi.exhaustive_topk(q, 100)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(_bits(d), _bits(od))


def test_me8q_diagnostics_against_oracle_at_size(ctx):
    from test_gpu_me8q import _oracle_case
    data, _c, _l, q, oix = _oracle_case(31, 24000, 320, 48, 3.0, spec="64:8,128:4,128:2")
    ix = capi_m8.M8Index.from_file(ctx, me8qx.parse(oix.image()))
    rng = np.random.default_rng(4)
    pq = rng.integers(0, q.shape[0], 500).astype(np.uint32)
    pid = rng.integers(0, data.shape[0], 500).astype(np.uint32)
    print("me8q diagnostics at size: estimate", flush=True)
    np.testing.assert_array_equal(_bits(ix.estimate(q, pq, pid)), _bits(oix.estimate(q, pq, pid)))
    print("me8q diagnostics at size: exhaustive k=256", flush=True)
    ids, d = ix.exhaustive_topk(q[:16], 256)
    oids, od = oix.exhaustive_topk(q[:16], 256)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(_bits(d), _bits(od))
    for k, nprobe in ((10, 4), (300, 12), (1024, 48)):
        print(f"me8q diagnostics at size: search_ivf k={k} nprobe={nprobe}", flush=True)
        ids, d, st = ix.search_ivf(q, k, nprobe)
        oids, od, ost = oix.search_ivf(q, k, nprobe)
        np.testing.assert_array_equal(ids, oids)
        np.testing.assert_array_equal(_bits(d), _bits(od))
        np.testing.assert_array_equal(st, ost)
    ix.close()


def test_module_diagnostics_against_reference(gold):
    # e8qlib's E8qIndex.estimate / probe_bound_w / exhaustive_topk and
    # Me8qIndex.estimate / exhaustive_topk / search_ivf on the reference's
    # own files: the drop-in Python surface (e8q_bindings.cpp:323-332,
    # me8q_bindings.cpp:304)
    import e8qlib
    name = "d128b2"
    data, _c, _l, queries = cases.mixture(name)
    idx = e8qlib.E8qIndex.load(os.path.join(G, f"{name}.e8ax"))
    pq, pid = cases.diag_pairs(name)
    est = np.array([idx.estimate(queries[q], int(i)) for q, i in zip(pq, pid)], np.float32)
    np.testing.assert_array_equal(_bits(est), _bits(gold[f"{name}_est"]))
    _same_f(idx.probe_bound_w(queries, data[pid], pq, pid, num_threads=4), gold[f"{name}_pbw"])
    ex = idx.exhaustive_topk(queries[3], 10)
    assert [e[1] for e in ex] == gold[f"{name}_k10_exids"][3].tolist()
    assert [np.float32(e[0]) for e in ex] == gold[f"{name}_k10_exd"][3].tolist()
    with pytest.raises(ValueError):
        idx.estimate(queries[0], 99999)
    name = "m256sp"
    _d, _c, _l, mq, _lam = cases.me8q_mixture(name)
    m = e8qlib.Me8qIndex.load(os.path.join(G, f"{name}.me8q"))
    pq, pid = cases.diag_pairs(name)
    est = np.array([m.estimate(mq[q], int(i)) for q, i in zip(pq, pid)], np.float32)
    np.testing.assert_array_equal(_bits(est), _bits(gold[f"{name}_est"]))
    ids, d = m.exhaustive_topk(mq, 100)
    np.testing.assert_array_equal(ids, gold[f"{name}_k100_exids"])
    np.testing.assert_array_equal(_bits(d), _bits(gold[f"{name}_k100_exd"]))
    nlist = cases.ME8Q_CASES[name][2]
    ids, d, st = m.search_ivf(mq, 10, 3, return_stats=True)
    np.testing.assert_array_equal(ids, gold[f"{name}_k10_p3_ivfids"])
    np.testing.assert_array_equal(_bits(d), _bits(gold[f"{name}_k10_p3_ivfd"]))
    assert st["candidates"] == int(gold[f"{name}_k10_p3_ivfstats"][:, 0].sum())
    with pytest.raises(IndexError):
        m.estimate(mq[0], 5000)
    assert nlist > 3
