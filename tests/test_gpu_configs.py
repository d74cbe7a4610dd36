"""The device path at configuration shapes, against fixtures the reference
library wrote (tests/golden/make_golden.py --only flush,c1,c3):

* survivor flush sizes 1, 7, 256 and 1024 (E8Q_FLUSH_EVERY: the reference
  stages up to 1024 survivors, src/index/e8q_internal.hpp:33), through the C
  ABI's flush_every and through the module's environment knob; beyond 1024
  the C ABI answers E8Q_EINVAL;
* the CPU-reference E8Q configuration at full size (100k x 128 i.i.d.
  N(0, 1), 256 random-row centroids, b = 4, k = 10, nprobe = 16): the device
  build's saved E8AX file section by section, and the search of 1000 queries
  (ids, distance bits, candidate / survivor counts);
* the mixed-rate ME8Q shape at D = 960 (SAQ DP plan at avg_bits 4 with the
  power-law variance, 256 lists): plan, the saved ME8QIDXX image, the plain
  searches bit for bit, and the batched form (every returned distance is that
  vector's exact estimator distance from the oracle's unpruned search).
"""
import json
import os
import tempfile

import numpy as np
import pytest

import capi
import cases
import e8ax
import oracle_py as orc
from tables import tables

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(G, "manifest.json")))
INF = float("inf")


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


def _z(k, alpha=1.0):
    return orc.lib().orc_ortho_quantile(alpha, max(4 * k, 32))


@pytest.mark.parametrize("name", cases.FLUSH_INDEXES)
def test_flush_sizes_c_abi(ctx, name):
    gold = np.load(os.path.join(G, "flush.npz"))
    a = e8ax.read(os.path.join(G, f"{name}.e8ax"))
    ix = capi.DeviceIndex(ctx, a, tables(a["bits"], a["padded"]))
    ix.upload()
    _d, _c, _l, queries = cases.mixture(name)
    for fe in cases.FLUSH_SIZES:
        for (k, nprobe, dis) in cases.flush_grid(name):
            tag = f"{name}_f{fe}_k{k}_p{nprobe}_d{dis}"
            print(f"flush size {tag}", flush=True)
            ids, d, st = ix.search(queries, k, nprobe, _z(k, a["alpha"]), disable_prune=bool(dis), flush_every=fe)
            np.testing.assert_array_equal(ids, gold[f"{tag}_ids"], err_msg=tag)
            np.testing.assert_array_equal(d.view(np.uint32), gold[f"{tag}_dists"].view(np.uint32), err_msg=tag)
            assert st == (int(gold[f"{tag}_stats"][:, 0].sum()), int(gold[f"{tag}_stats"][:, 1].sum())), tag


def test_flush_size_env_and_cap(ctx, monkeypatch):
    import e8qlib
    gold = np.load(os.path.join(G, "flush.npz"))
    _d, _c, _l, queries = cases.mixture("d128b4")
    idx = e8qlib.E8qIndex.load(os.path.join(G, "d128b4.e8ax"))
    k, nprobe, dis = cases.flush_grid("d128b4")[1]
    for fe in (7, 1024):
        print(f"flush size {fe} through E8Q_FLUSH_EVERY", flush=True)
        monkeypatch.setenv("E8Q_FLUSH_EVERY", str(fe))
        ids, d = idx.search(queries, k, nprobe)
        tag = f"d128b4_f{fe}_k{k}_p{nprobe}_d{dis}"
        np.testing.assert_array_equal(ids, gold[f"{tag}_ids"])
        np.testing.assert_array_equal(d.view(np.uint32), gold[f"{tag}_dists"].view(np.uint32))
    monkeypatch.delenv("E8Q_FLUSH_EVERY")
    a = e8ax.read(os.path.join(G, "d128b4.e8ax"))
    ix = capi.DeviceIndex(ctx, a, tables(a["bits"], a["padded"]))
    ix.upload()
    with pytest.raises(RuntimeError, match="flush_every must be <= 1024"):
        ix.search(queries, 10, 4, _z(10), flush_every=1025)


def test_c1_iid_full_size(tmp_path):
    import e8qlib
    base, cent, queries = cases.c1_iid()
    gold = np.load(os.path.join(G, "c1_iid.npz"))
    lab = gold["labels"].astype(np.uint32)
    c = cases.C1
    print("c1: device build of 100k x 128", flush=True)
    idx = e8qlib.E8qIndex(c["dim"], c["bits"], c["seed"])
    idx.build(base, cent, lab)
    path = str(tmp_path / "c1.e8ax")
    idx.save(path)
    assert cases.e8ax_digests(e8ax.read(path)) == MAN["c1"]["image"]
    print("c1: search of 1000 queries, k=10, nprobe=16", flush=True)
    ids, d, s = idx.search(queries, c["k"], c["nprobe"], return_stats=True)
    np.testing.assert_array_equal(ids, gold["ids"])
    np.testing.assert_array_equal(d.view(np.uint32), gold["dists"].view(np.uint32))
    assert s["candidates"] == int(gold["stats"][:, 0].sum())
    assert s["survivors"] == int(gold["stats"][:, 1].sum())


@pytest.fixture(scope="module")
def c3_index():
    import e8qlib
    data, cent, queries, lam = cases.c3_shape()
    gold = np.load(os.path.join(G, "c3_shape.npz"))
    lab = gold["labels"].astype(np.uint32)
    c = cases.C3
    print("c3: device build of 20k x 960", flush=True)
    idx = e8qlib.Me8qIndex(c["dim"], c["avg_bits"], c["seed"])
    idx.build(data, cent, lab, variance=lam)
    oix = orc.OracleMe8q.build(data, cent, lab, c["avg_bits"], c["seed"], 0.0, "", lam, True)
    return idx, oix, gold, queries


def test_c3_shape_plan_and_image(c3_index):
    idx, _oix, gold, _q = c3_index
    assert [tuple(p) for p in idx.plan] == [tuple(int(v) for v in r) for r in gold["plan"].tolist()]
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "c3.me8q")
        idx.save(path)
        assert cases.digest(np.fromfile(path, np.uint8)) == MAN["c3"]["image_sha256"]


@pytest.mark.parametrize("k,nprobe", cases.c3_grid())
def test_c3_shape_plain_search(c3_index, k, nprobe):
    idx, _oix, gold, queries = c3_index
    print(f"c3: plain search k={k} nprobe={nprobe}", flush=True)
    ids, d, s = idx.search(queries, k, nprobe, return_stats=True)
    tag = f"k{k}_p{nprobe}"
    np.testing.assert_array_equal(ids, gold[f"{tag}_ids"])
    np.testing.assert_array_equal(d.view(np.uint32), gold[f"{tag}_dists"].view(np.uint32))
    assert s["candidates"] == int(gold[f"{tag}_stats"][:, 0].sum())
    assert s["survivors"] == int(gold[f"{tag}_stats"][:, 1].sum())
    assert s["bits_scored"] == int(gold[f"{tag}_stats"][:, 2].sum())


@pytest.mark.parametrize("k,nprobe", cases.c3_grid())
def test_c3_shape_batched_search(c3_index, k, nprobe):
    idx, oix, gold, queries = c3_index
    print(f"c3: batched search k={k} nprobe={nprobe}", flush=True)
    bi, bd, s = idx.search(queries, k, nprobe, return_stats=True, mode="batched")
    assert s["form"] == "batched"
    ui, ud, _ = oix.search(queries, 2048, nprobe, z=INF)
    tag = f"k{k}_p{nprobe}"
    pi, pd = gold[f"{tag}_ids"], gold[f"{tag}_dists"]
    found = 0
    for i in range(queries.shape[0]):
        exact = {int(a): np.float32(b) for a, b in zip(ui[i], ud[i]) if a != 0xFFFFFFFF}
        for a, b in zip(bi[i], bd[i]):
            if a == 0xFFFFFFFF:
                continue
            assert int(a) in exact, f"query {i}: id {a} is not among the nearest probed vectors"
            assert abs(int(np.float32(b).view(np.int32)) - int(exact[int(a)].view(np.int32))) <= 1
        kb, kp = bd[i][bd[i] < INF], pd[i][pd[i] < INF]
        if kb.size == kp.size and kb.size > 0:
            assert kb[-1] <= kp[-1] + 1e-6 * max(1.0, abs(float(kp[-1])))
        found += len(set(map(int, bi[i])) & set(map(int, pi[i])))
    assert found / float(queries.shape[0] * k) >= 0.97
