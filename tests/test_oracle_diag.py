"""The oracle's diagnostics against fixtures written by the reference library
(tests/golden/make_golden.py --only diag, through the reference's own
E8qIndex::Estimate / ProbeBoundW / ExhaustiveTopK, e8q_oracle.cpp:14-219,
and Me8qIndex::Estimate / ExhaustiveTopK / SearchIVF, me8q_measure.cpp:41-67
and me8q_search.cpp:136-266), on every E8Q and ME8Q fixture index.

Everything is compared exactly: float and double bits (NaN where the
reference returns NaN: the row that sits on its centroid has no residual),
id lists in order (ties included: the ME8Q lists keep the reference's
insertion order among equal distances), SearchIVF's counters.  The E8Q
fixtures cover b = 1, 2, 4, 8 and 9, where ProbeBoundW decodes with the
reference's 8-bit digit form (its b = 1 and b = 9 quirks included).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402
import e8ax  # noqa: E402
import oracle_py as orc  # noqa: E402

G = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(G, "diag.npz"))


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_e8q_diagnostics(gold, name):
    data, _c, _l, queries = cases.mixture(name)
    ix = orc.OracleIndex.from_e8ax(e8ax.read(os.path.join(G, f"{name}.e8ax")))
    pq, pid = cases.diag_pairs(name)
    est = ix.estimate(queries, pq, pid)
    np.testing.assert_array_equal(_bits(est), _bits(gold[f"{name}_est"]))
    pbw = ix.probe_bound_w(queries, data[pid], pq, pid)
    want = gold[f"{name}_pbw"]
    assert np.isnan(want[-2:]).all()   # the rows on their centroid
    np.testing.assert_array_equal(np.isnan(pbw), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_array_equal(_bits(pbw[ok]), _bits(want[ok]))
    for k in cases.diag_k(name):
        ids, d = ix.exhaustive_topk(queries, k)
        np.testing.assert_array_equal(ids, gold[f"{name}_k{k}_exids"])
        np.testing.assert_array_equal(_bits(d), _bits(gold[f"{name}_k{k}_exd"]))


def test_e8q_estimate_unknown_id():
    ix = orc.OracleIndex.from_e8ax(e8ax.read(os.path.join(G, "d128b4.e8ax")))
    _d, _c, _l, queries = cases.mixture("d128b4")
    with pytest.raises(KeyError):
        ix.estimate(queries, [0], [3000])


@pytest.fixture(scope="module", params=list(cases.ME8Q_CASES))
def m8(request):
    name = request.param
    n, dim, nlist, avg, seed, alpha, spec, pv, de, _e, _nq = cases.ME8Q_CASES[name]
    data, cent, lab, queries, lam = cases.me8q_mixture(name)
    ix = orc.OracleMe8q.build(data, cent, lab, avg, seed, alpha, spec, lam if pv else None, de)
    return name, ix, queries


def test_me8q_diagnostics(gold, m8):
    name, ix, queries = m8
    pq, pid = cases.diag_pairs(name)
    np.testing.assert_array_equal(_bits(ix.estimate(queries, pq, pid)), _bits(gold[f"{name}_est"]))
    for k in sorted({k for k, _ in cases.diag_ivf_grid(name)}):
        ids, d = ix.exhaustive_topk(queries, k)
        np.testing.assert_array_equal(ids, gold[f"{name}_k{k}_exids"])
        np.testing.assert_array_equal(_bits(d), _bits(gold[f"{name}_k{k}_exd"]))
    for (k, nprobe) in cases.diag_ivf_grid(name):
        ids, d, st = ix.search_ivf(queries, k, nprobe)
        tag = f"{name}_k{k}_p{nprobe}"
        np.testing.assert_array_equal(ids, gold[f"{tag}_ivfids"])
        np.testing.assert_array_equal(_bits(d), _bits(gold[f"{tag}_ivfd"]))
        np.testing.assert_array_equal(st, gold[f"{tag}_ivfstats"])


def test_me8q_ties_kept_in_insertion_order(gold):
    # m64drop's 400-wide lists hold runs of equal distances: the reference's
    # PushCandidate puts a later equal item first and rejects one equal to a
    # full list's worst, so the order is not a sort by (distance, id)
    d = gold["m64drop_k400_exd"]
    ids = gold["m64drop_k400_exids"]
    ties = 0
    for row_d, row_i in zip(d, ids):
        fin = row_d < np.inf
        for a in np.unique(row_d[fin]):
            sel = row_i[fin][row_d[fin] == a]
            if sel.size > 1:
                ties += 1
                assert not np.all(np.diff(sel.astype(np.int64)) > 0)
    assert ties > 0
