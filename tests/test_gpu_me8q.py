"""ME8Q on the device, through the C ABI (include/me8q_gpu.h), against the
reference library's own outputs (tests/golden/*.me8q and *_me8q.npz, written
by gen_golden.cpp) and against the oracle (oracle/me8q_oracle.cpp, itself
pinned to those fixtures by test_oracle_me8q.py).  Everything is compared
exactly: code bytes, plane / refine / bank bytes, float bits of the fits,
rotated queries and centroids, LUT bytes, probe lists, result ids, distance
bits and the per-query candidate / survivor / bit counters."""
import os

import numpy as np
import pytest

import capi
import capi_m8
import cases
import me8qx
import oracle_py as orc

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = list(cases.ME8Q_CASES)


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


def _z(k, alpha):
    return orc.lib().orc_ortho_quantile(alpha if alpha > 0 else 1.0, max(4 * k, 32))


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a.view(np.uint64)


def _oracle(name):
    n, dim, nlist, avg, seed, alpha, spec, pv, de, _e, _nq = cases.ME8Q_CASES[name]
    data, cent, lab, queries, lam = cases.me8q_mixture(name)
    return orc.OracleMe8q.build(data, cent, lab, avg, seed, alpha, spec, lam if pv else None, de)


def _check_fast(ix, gold, nlist):
    crot, cn2 = ix.centroids_rot()
    np.testing.assert_array_equal(_bits(crot), _bits(gold["crot"]))
    np.testing.assert_array_equal(_bits(cn2), _bits(gold["cn2"]))
    off = gold["off"]
    for c in range(nlist):
        plane, refine, gain, vpc, rcn2 = ix.fast(c)
        np.testing.assert_array_equal(refine, gold["refine"][off[c, 1]:off[c + 1, 1]], err_msg=f"refine list {c}")
        np.testing.assert_array_equal(_bits(gain), _bits(gold["gain"][off[c, 2]:off[c + 1, 2]]), err_msg=f"gain {c}")
        np.testing.assert_array_equal(_bits(vpc), _bits(gold["vpc"][off[c, 2]:off[c + 1, 2]]), err_msg=f"vpc {c}")
        np.testing.assert_array_equal(_bits(rcn2), _bits(gold["rcn2"][off[c, 3]:off[c + 1, 3]]), err_msg=f"rcn2 {c}")
        np.testing.assert_array_equal(plane, gold["plane"][off[c, 0]:off[c + 1, 0]], err_msg=f"plane list {c}")


@pytest.mark.parametrize("name", NAMES)
def test_loaded_index_layout(ctx, name):
    print(f"me8q layout {name}", flush=True)
    f = me8qx.read(os.path.join(G, f"{name}.me8q"))
    gold = np.load(os.path.join(G, f"{name}_me8q.npz"))
    ix = capi_m8.M8Index.from_file(ctx, f)
    _check_fast(ix, gold, f["nlist"])
    ids, cent, codes = ix.download()
    np.testing.assert_array_equal(ids, f["ids"])
    np.testing.assert_array_equal(codes, f["codes"])


@pytest.mark.parametrize("name", NAMES)
def test_search_matches_reference(ctx, name):
    f = me8qx.read(os.path.join(G, f"{name}.me8q"))
    gold = np.load(os.path.join(G, f"{name}_me8q.npz"))
    queries = gold["queries"]
    ix = capi_m8.M8Index.from_file(ctx, f)
    oix = _oracle(name)
    for (k, nprobe) in cases.me8q_search_grid(name):
        tag = f"k{k}_p{nprobe}"
        print(f"me8q search {name} {tag}", flush=True)
        rq, lut, md, probe, pd = ix.query(queries, nprobe)
        np.testing.assert_array_equal(_bits(rq), _bits(gold[f"{tag}_rotq"]))
        np.testing.assert_array_equal(probe, gold[f"{tag}_probe"])
        np.testing.assert_array_equal(_bits(pd), _bits(gold[f"{tag}_pdist"]))
        for i in range(queries.shape[0]):
            _rq, olut, omd, _pr, _pd = oix.query(queries[i], nprobe)
            np.testing.assert_array_equal(lut[i], olut, err_msg=f"LUT query {i}")
            np.testing.assert_array_equal(_bits(md[i]), _bits(omd))
        ids, d, st, tot = ix.search(queries, k, nprobe, _z(k, f["prune_alpha"]))
        np.testing.assert_array_equal(ids, gold[f"{tag}_ids"], err_msg=tag)
        np.testing.assert_array_equal(_bits(d), _bits(gold[f"{tag}_dists"]), err_msg=tag)
        np.testing.assert_array_equal(st, gold[f"{tag}_stats"], err_msg=tag)
        assert tot.candidates == int(gold[f"{tag}_stats"][:, 0].sum())


@pytest.mark.parametrize("name", NAMES)
def test_device_encode_matches_reference_file(ctx, name):
    print(f"me8q encode {name}", flush=True)
    n, dim, nlist, *_ = cases.ME8Q_CASES[name]
    data, cent, lab, _q, _lam = cases.me8q_mixture(name)
    f = me8qx.read(os.path.join(G, f"{name}.me8q"))
    gold = np.load(os.path.join(G, f"{name}_me8q.npz"))
    ix = capi_m8.M8Index(ctx, n, dim, nlist, f["plan"], f["signs"], capi_m8.list_begin(lab, nlist), f["drop_energy"])
    cp = np.zeros((nlist, f["padded"]), np.float32)
    cp[:, :dim] = cent
    ix.encode(data, cp, lab, capi_m8.ranks(lab, nlist))
    ids, cd, codes = ix.download()
    np.testing.assert_array_equal(ids, f["ids"])
    np.testing.assert_array_equal(_bits(cd), _bits(f["centroids"]))
    bad = np.flatnonzero(codes != f["codes"])
    assert bad.size == 0, f"{bad.size} code bytes differ, first at {bad[:8]}"
    _check_fast(ix, gold, nlist)


def _oracle_case(seed, n, dim, nlist, avg, spec="", alpha=0.0, sizes=None):
    rng = np.random.default_rng(seed)
    lam = (np.arange(1, dim + 1, dtype=np.float64) ** -1.0).astype(np.float32)
    sd = np.sqrt(lam).astype(np.float32)
    cent = (rng.standard_normal((nlist, dim)).astype(np.float32) * 3.0 * sd).astype(np.float32)
    if sizes is None:
        lab = rng.integers(0, nlist, n).astype(np.uint32)
    else:
        lab = np.repeat(np.arange(nlist, dtype=np.uint32), sizes)
        rng.shuffle(lab)
        n = lab.size
    data = (cent[lab] + rng.standard_normal((n, dim)).astype(np.float32) * sd).astype(np.float32)
    q = (cent[rng.integers(0, nlist, 64)] + rng.standard_normal((64, dim)).astype(np.float32) * sd).astype(np.float32)
    oix = orc.OracleMe8q.build(data, cent, lab, avg, seed, alpha, spec)
    return data, cent, lab, q, oix


def test_larger_index_against_oracle(ctx):
    data, cent, lab, q, oix = _oracle_case(31, 24000, 320, 48, 3.0)
    f = me8qx.parse(oix.image())
    print(f"me8q larger: plan {f['plan']}", flush=True)
    ix = capi_m8.M8Index.from_file(ctx, f)
    for k, nprobe in ((10, 8), (100, 24), (1, 48), (300, 6)):
        print(f"me8q larger k={k} nprobe={nprobe}", flush=True)
        ids, d, st, _ = ix.search(q, k, nprobe, _z(k, 0.0))
        oids, od, ost = oix.search(q, k, nprobe)
        np.testing.assert_array_equal(ids, oids)
        np.testing.assert_array_equal(_bits(d), _bits(od))
        np.testing.assert_array_equal(st, ost)
    # the device encode of the same rows reproduces the oracle's image
    ix2 = capi_m8.M8Index(ctx, f["n"], f["dim"], f["nlist"], f["plan"], f["signs"], f["list_begin"], f["drop_energy"])
    cp = np.zeros((f["nlist"], f["padded"]), np.float32)
    cp[:, :320] = cent
    ix2.encode(data, cp, lab, capi_m8.ranks(lab, f["nlist"]))
    _ids, _c, codes = ix2.download()
    np.testing.assert_array_equal(codes, f["codes"])


def test_two_stage_probe_selection_against_oracle(ctx):
    # 512 lists and nprobe <= 64 take the two-stage selection (approximate
    # MFMA route distances, the reference's chain order for the candidates
    # their error bound leaves): probe lists and distances must be the
    # oracle's bit for bit, and so must the search.
    data, cent, lab, q, oix = _oracle_case(7, 12000, 128, 512, 4.0)
    f = me8qx.parse(oix.image())
    ix = capi_m8.M8Index.from_file(ctx, f)
    for nprobe in (1, 16, 64):
        print(f"me8q two-stage route nprobe={nprobe}", flush=True)
        _rq, _lut, _md, probe, pd = ix.query(q, nprobe)
        for i in range(q.shape[0]):
            _r, _l, _m, opr, opd = oix.query(q[i], nprobe)
            np.testing.assert_array_equal(probe[i], np.asarray(opr).reshape(-1), err_msg=f"query {i}")
            np.testing.assert_array_equal(_bits(pd[i]), _bits(np.asarray(opd, np.float32).reshape(-1)))
    ids, d, st, _ = ix.search(q, 10, 16, _z(10, 0.0))
    oids, od, ost = oix.search(q, 10, 16)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(_bits(d), _bits(od))


def test_uneven_nprobe_and_skewed_lists(ctx):
    sizes = np.array([3000, 5, 0, 700, 1, 64, 33, 2500, 31, 900], np.int64)
    data, cent, lab, q, oix = _oracle_case(77, 0, 256, 10, 4.0, spec="64:6,128:4,64:2", sizes=sizes)
    f = me8qx.parse(oix.image())
    ix = capi_m8.M8Index.from_file(ctx, f)
    npq = np.random.default_rng(5).choice(np.array([1, 2, 3, 7, 10], np.uint32), size=q.shape[0])
    print("me8q uneven nprobe", flush=True)
    ids, d, st, _ = ix.search(q, 20, 10, _z(20, 0.0), nprobe_per_query=npq)
    for i in range(q.shape[0]):
        oids, od, ost = oix.search(q[i:i + 1], 20, int(npq[i]))
        np.testing.assert_array_equal(ids[i], oids[0])
        np.testing.assert_array_equal(_bits(d[i]), _bits(od[0]))
        np.testing.assert_array_equal(st[i], ost[0])


def test_python_api(tmp_path):
    import e8qlib
    for name in ("m400a4", "m256sp", "m64drop"):
        print(f"me8q python api {name}", flush=True)
        n, dim, nlist, avg, seed, alpha, spec, pv, de, _e, _nq = cases.ME8Q_CASES[name]
        data, cent, lab, queries, lam = cases.me8q_mixture(name)
        gold = np.load(os.path.join(G, f"{name}_me8q.npz"))
        idx = e8qlib.Me8qIndex(dim, avg, seed)
        idx.build(data, cent, lab, variance=lam if pv else None, drop_energy=de, prune_alpha=alpha, plan_spec=spec)
        assert [tuple(s) for s in idx.plan] == [tuple(r) for r in gold["plan"].tolist()]
        path = str(tmp_path / f"{name}.me8q")
        idx.save(path)
        ref = open(os.path.join(G, f"{name}.me8q"), "rb").read()
        assert open(path, "rb").read() == ref
        k, nprobe = cases.me8q_search_grid(name)[2]
        tag = f"k{k}_p{nprobe}"
        for ix in (idx, e8qlib.Me8qIndex.load(os.path.join(G, f"{name}.me8q"))):
            ids, d, s = ix.search(queries, k, nprobe, return_stats=True)
            np.testing.assert_array_equal(ids, gold[f"{tag}_ids"])
            np.testing.assert_array_equal(_bits(d), _bits(gold[f"{tag}_dists"]))
            assert s["candidates"] == int(gold[f"{tag}_stats"][:, 0].sum())
            assert s["survivors"] == int(gold[f"{tag}_stats"][:, 1].sum())
