"""The oracle against reference-written fixtures at configuration shapes
(tests/golden/make_golden.py --only flush,c1,c3):

* survivor flush sizes 1, 7, 256 and 1024 (E8Q_FLUSH_EVERY, read by the
  reference's FlushEvery, src/index/e8q_scan_common.hpp:66-73);
* the CPU-reference E8Q configuration at full size (100k x 128 i.i.d.
  N(0, 1), 256 random-row centroids, k = 10, nprobe = 16): every array of
  the index image (by SHA-256 of each section) and the search of 1000 queries;
* the mixed-rate ME8Q shape at D = 960 (DP plan at avg_bits 4 with the
  power-law variance, 256 lists, k = 100 / 10, nprobe 32 / 64): the plan, the
  whole ME8QIDXX image (SHA-256) and the searches;
* the oracle's unpruned ME8Q search (lambda = +inf), used as the yardstick of
  the batched form: it scores every probed vector, and the pruned search's
  survivors carry the same distances.
"""
import json
import os

import numpy as np
import pytest

import cases
import oracle_py as orc

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(G, "manifest.json")))
INF = float("inf")


def test_manifest_inputs_unchanged():
    base, cent, queries = cases.c1_iid()
    assert cases.digest(base, cent, queries) == MAN["c1"]["inputs"]
    data, cent3, q3, lam = cases.c3_shape()
    assert cases.digest(data, cent3, q3, lam) == MAN["c3"]["inputs"]
    assert MAN["flush"]["sizes"] == cases.FLUSH_SIZES


@pytest.mark.parametrize("name", cases.FLUSH_INDEXES)
@pytest.mark.parametrize("fe", cases.FLUSH_SIZES)
def test_flush_sizes(monkeypatch, name, fe):
    import e8ax
    gold = np.load(os.path.join(G, "flush.npz"))
    ix = orc.OracleIndex.from_e8ax(e8ax.read(os.path.join(G, f"{name}.e8ax")))
    _d, _c, _l, queries = cases.mixture(name)
    monkeypatch.setenv("E8Q_FLUSH_EVERY", str(fe))
    for (k, nprobe, dis) in cases.flush_grid(name):
        tag = f"{name}_f{fe}_k{k}_p{nprobe}_d{dis}"
        ids, d, st = ix.search(queries, k, nprobe, disable_prune=bool(dis))
        np.testing.assert_array_equal(ids, gold[f"{tag}_ids"], err_msg=tag)
        np.testing.assert_array_equal(d.view(np.uint32), gold[f"{tag}_dists"].view(np.uint32), err_msg=tag)
        assert int(st[0]) == int(gold[f"{tag}_stats"][:, 0].sum())
        assert int(st[1]) == int(gold[f"{tag}_stats"][:, 1].sum())


def test_flush_size_changes_results():
    # the knob is not inert on these fixtures: the survivor counts differ
    gold = np.load(os.path.join(G, "flush.npz"))
    k, nprobe, dis = cases.flush_grid("d128b4")[1]
    surv = {fe: int(gold[f"d128b4_f{fe}_k{k}_p{nprobe}_d{dis}_stats"][:, 1].sum()) for fe in cases.FLUSH_SIZES}
    assert len(set(surv.values())) > 1, surv


def test_c1_iid_full_size():
    base, cent, queries = cases.c1_iid()
    gold = np.load(os.path.join(G, "c1_iid.npz"))
    lab = gold["labels"].astype(np.uint32)
    c = cases.C1
    ix = orc.OracleIndex.build(base, cent, lab, c["bits"], c["seed"])
    got = cases.e8ax_digests(ix.arrays())
    assert got == MAN["c1"]["image"]
    ids, d, st = ix.search(queries, c["k"], c["nprobe"], threads=8)
    np.testing.assert_array_equal(ids, gold["ids"])
    np.testing.assert_array_equal(d.view(np.uint32), gold["dists"].view(np.uint32))
    assert int(st[0]) == int(gold["stats"][:, 0].sum())
    assert int(st[1]) == int(gold["stats"][:, 1].sum())


@pytest.fixture(scope="module")
def c3_oracle():
    data, cent, queries, lam = cases.c3_shape()
    gold = np.load(os.path.join(G, "c3_shape.npz"))
    lab = gold["labels"].astype(np.uint32)
    c = cases.C3
    ix = orc.OracleMe8q.build(data, cent, lab, c["avg_bits"], c["seed"], 0.0, "", lam, True)
    return ix, gold, queries


def test_c3_shape_plan_and_image(c3_oracle):
    ix, gold, _q = c3_oracle
    assert ix.plan == [tuple(int(v) for v in r) for r in gold["plan"].tolist()]
    assert cases.digest(ix.image()) == MAN["c3"]["image_sha256"]


@pytest.mark.parametrize("k,nprobe", cases.c3_grid())
def test_c3_shape_search(c3_oracle, k, nprobe):
    ix, gold, queries = c3_oracle
    ids, d, st = ix.search(queries, k, nprobe)
    tag = f"k{k}_p{nprobe}"
    np.testing.assert_array_equal(ids, gold[f"{tag}_ids"])
    np.testing.assert_array_equal(d.view(np.uint32), gold[f"{tag}_dists"].view(np.uint32))
    np.testing.assert_array_equal(st, gold[f"{tag}_stats"])


@pytest.mark.parametrize("k,nprobe", [(100, 32), (10, 64)])
def test_me8q_unpruned_search(c3_oracle, k, nprobe):
    # lambda = +inf prunes nothing: every probed vector is a survivor, the
    # list is sorted, the pruned search's k-th distance is never below it,
    # and each vector the pruned search returns carries (to the ulp the
    # cascade's group order allows) the distance the unpruned search gives it.
    ix, gold, queries = c3_oracle
    big = 4000
    ui, ud, ust = ix.search(queries, big, nprobe, z=INF)
    np.testing.assert_array_equal(ust[:, 1], ust[:, 0])   # survivors == candidates
    assert np.all(ust[:, 4] == 0)
    pi, pd, _ = ix.search(queries, k, nprobe)
    for i in range(queries.shape[0]):
        valid = ui[i] != 0xFFFFFFFF
        assert int(valid.sum()) == min(big, int(ust[i, 0]))
        assert np.all(np.diff(ud[i][valid]) >= 0)
        exact = dict(zip(ui[i][valid].tolist(), ud[i][valid].tolist()))
        for a, b in zip(pi[i], pd[i]):
            if a == 0xFFFFFFFF:
                continue
            e = np.float32(exact[int(a)])
            assert abs(int(np.float32(b).view(np.int32)) - int(e.view(np.int32))) <= 1
        kk = min(k, int(valid.sum()))
        assert ud[i][kk - 1] <= pd[i][kk - 1]
