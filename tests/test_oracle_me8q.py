"""The ME8Q oracle (oracle/me8q_oracle.cpp) against fixtures written by the
reference library (tests/golden/gen_golden.cpp, make_golden.py): plan, whole
ME8QIDXX image, fast layout (plane/refine bytes, u8 banks, gain/vpc, per-run
centroid energies), rotated centroids, rotated queries, probe lists and the
search results -- ids, distance bits and the candidate/survivor/bit
counters -- all compared exactly."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402
import oracle_py as orc  # noqa: E402

G = os.path.join(HERE, "golden")
NAMES = list(cases.ME8Q_CASES)


@pytest.fixture(scope="module", params=NAMES)
def built(request):
    name = request.param
    n, dim, nlist, avg, seed, alpha, spec, pv, de, _e, _nq = cases.ME8Q_CASES[name]
    data, cent, lab, queries, lam = cases.me8q_mixture(name)
    ix = orc.OracleMe8q.build(data, cent, lab, avg, seed, alpha, spec, lam if pv else None, de)
    gold = np.load(os.path.join(G, f"{name}_me8q.npz"))
    return name, ix, gold, queries


def test_plan_and_image(built):
    name, ix, gold, _ = built
    assert [tuple(r) for r in gold["plan"].tolist()] == ix.plan
    assert int(gold["routedim"][0]) == ix.route_dim
    ref = np.fromfile(os.path.join(G, f"{name}.me8q"), np.uint8)
    img = ix.image()
    assert img.size == ref.size
    bad = np.flatnonzero(img != ref)
    assert bad.size == 0, f"{bad.size} image bytes differ, first at {bad[:8]}"


def test_fast_layout(built):
    name, ix, gold, _ = built
    crot, cn2 = ix.centroids()
    np.testing.assert_array_equal(crot.view(np.uint32), gold["crot"].view(np.uint32))
    np.testing.assert_array_equal(cn2.view(np.uint32), gold["cn2"].view(np.uint32))
    off = gold["off"]
    for c in range(ix.nlist):
        plane, refine, gain, vpc, rcn2 = ix.fast(c)
        np.testing.assert_array_equal(refine, gold["refine"][off[c, 1]:off[c + 1, 1]], err_msg=f"refine list {c}")
        np.testing.assert_array_equal(gain.view(np.uint32), gold["gain"][off[c, 2]:off[c + 1, 2]].view(np.uint32),
                                      err_msg=f"gain list {c}")
        np.testing.assert_array_equal(vpc.view(np.uint32), gold["vpc"][off[c, 2]:off[c + 1, 2]].view(np.uint32),
                                      err_msg=f"vpc list {c}")
        np.testing.assert_array_equal(rcn2.view(np.uint64), gold["rcn2"][off[c, 3]:off[c + 1, 3]].view(np.uint64))
        np.testing.assert_array_equal(plane, gold["plane"][off[c, 0]:off[c + 1, 0]], err_msg=f"plane list {c}")


def test_query_side(built):
    name, ix, gold, queries = built
    for (k, nprobe) in cases.me8q_search_grid(name):
        tag = f"k{k}_p{nprobe}"
        for i in range(queries.shape[0]):
            rq, _lut, _md, probe, pd = ix.query(queries[i], nprobe)
            np.testing.assert_array_equal(rq.view(np.uint32), gold[f"{tag}_rotq"][i].view(np.uint32))
            np.testing.assert_array_equal(probe, gold[f"{tag}_probe"][i])
            np.testing.assert_array_equal(pd.view(np.uint32), gold[f"{tag}_pdist"][i].view(np.uint32))


def test_search(built):
    name, ix, gold, queries = built
    for (k, nprobe) in cases.me8q_search_grid(name):
        tag = f"k{k}_p{nprobe}"
        ids, d, st = ix.search(queries, k, nprobe)
        np.testing.assert_array_equal(ids, gold[f"{tag}_ids"], err_msg=tag)
        np.testing.assert_array_equal(d.view(np.uint32), gold[f"{tag}_dists"].view(np.uint32), err_msg=tag)
        np.testing.assert_array_equal(st, gold[f"{tag}_stats"], err_msg=tag)


def test_fma_order_probe(tmp_path):
    """GCC fuses the first product of a sum of products (the order the oracle
    uses for the class-mean LUT and the ZPR fit's |R_x|^2), checked by
    compiling those expression shapes with the reference's flags."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = os.path.join(os.path.dirname(HERE), "oracle", "fma_probe.cpp")
    exe = str(tmp_path / "probe")
    subprocess.run(["g++", "-std=c++17", "-O3", "-march=x86-64-v3", "-fno-math-errno", src, "-o", exe], check=True)
    n, lf, ls, rf, rs = (int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                                         check=True).stdout.split())
    assert lf == n and rf == n
    assert ls < n and rs < n      # the other order is distinguishable
