"""Pins the oracle (oracle/e8q_oracle.cpp) to the reference's golden vectors.

Every fixture under tests/golden/ was written by the reference e8qlib library
itself (tests/golden/make_golden.py); comparisons are bit-exact except where
noted.  Plus closed-form / brute-force checks that do not depend on the
reference at all, chosen so a dropped term, a wrong sign or a transposed
operand in the oracle fails one of them.
"""
import itertools
import json
import os

import numpy as np
import pytest

import cases
import e8ax
import oracle_py as orc

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gl(name):
    return np.load(os.path.join(G, name))


# ---------------------------------------------------------------- level tables
def test_levels_bit_exact():
    flat = gl("levels.npy")
    off = 0
    for b in range(1, 14):
        for br in (0, 1):
            ref = flat[off: off + (1 << b)]
            off += 1 << b
            np.testing.assert_array_equal(orc.branch_levels(b, br), ref, err_msg=f"b={b} br={br}")


def test_fixed_scale_and_class_means():
    sc = gl("fixed_scale.npy")
    wc = gl("class_means.npy")
    for b in range(1, 14):
        assert orc.lib().orc_fixed_scale(b) == sc[b]
        np.testing.assert_array_equal(orc.class_means(b), wc[b])


def test_level_identities():
    # Branch reflection L1[q-1-d] = -L0[d] and class-mean reflection
    # w[1] = -w[0] (reference tests/unit/e8q_test.cpp:35-67); sortedness in
    # quantile order; Phi^-1 symmetry.
    for b in range(1, 14):
        q = 1 << b
        l0, l1 = orc.branch_levels(b, 0), orc.branch_levels(b, 1)
        assert np.max(np.abs(l1[::-1] + l0)) < 1e-9
        w = orc.class_means(b)
        assert abs(w[0] + w[1]) < 1e-9 and w[0] > 0  # top-bit class 1 = negative quantiles
        # Digit order along increasing quantile: q/2, ..., q-1, 0, ..., q/2-1.
        order = np.r_[np.arange(q // 2, q), np.arange(0, q // 2)]
        assert np.all(np.diff(l0[order]) > 0)
    # Level values against a closed-form normal quantile (scipy), b = 4.
    from scipy.stats import norm
    b, q = 4, 16
    t = np.arange(-q // 2, q // 2)
    ref = np.sqrt(1.25) * norm.ppf(0.5 + (t + 0.25) / q)
    got = orc.branch_levels(b, 0)[t % q]
    assert np.max(np.abs(got - ref)) < 2e-8   # Acklam: |rel err| < 1.15e-9 in p


def test_ortho_quantile():
    from scipy.stats import norm
    for alpha, keff in [(1.0, 400), (1.0, 40), (1e-6, 40), (0.01, 32), (50.0, 32)]:
        tail = min(alpha / keff, 0.49)
        want = max(norm.ppf(1 - tail), 1.0)
        got = orc.lib().orc_ortho_quantile(alpha, keff)
        # 12 Newton steps on erf from the asymptotic start: ~4e-5 relative
        # in the deep tail (tail 2.5e-8), far tighter elsewhere.
        assert abs(got - want) < 1e-4 * max(1.0, want)


# ---------------------------------------------------------------- rotator
@pytest.mark.parametrize("dim,seed", cases.ROTATE_CASES)
def test_rotator_bit_exact(dim, seed):
    ref = gl(f"rotate_{dim}.npz")
    np.testing.assert_array_equal(orc.rotator_signs(dim, seed), ref["signs"])
    np.testing.assert_array_equal(orc.rotate(dim, seed, cases.rotate_input(dim)), ref["out"])


@pytest.mark.parametrize("dim", [64, 128, 100, 768])
def test_rotator_is_orthogonal(dim):
    # Norm preservation and linearity (an orthogonal map of the zero-padded
    # input); catches a wrong scale, a dropped sign round or a bad window.
    rng = np.random.default_rng(dim)
    x = rng.standard_normal((4, dim)).astype(np.float32)
    y = orc.rotate(dim, 3, x).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(y, axis=1), np.linalg.norm(x, axis=1), rtol=1e-5)
    E = orc.rotate(dim, 3, np.eye(dim, dtype=np.float32)).astype(np.float64)
    G_ = E @ E.T
    np.testing.assert_allclose(G_, np.eye(dim), atol=2e-6)


# ---------------------------------------------------------------- encoder
def test_encoder_bit_exact():
    ref = gl("encode.npz")
    u = cases.encode_input()
    for b in range(1, 14):
        dg, br = orc.encode(u, b)
        np.testing.assert_array_equal(dg, ref[f"digits_b{b}"], err_msg=f"b={b}")
        np.testing.assert_array_equal(br, ref[f"branch_b{b}"], err_msg=f"b={b}")


def _levels_for(bits):
    return orc.branch_levels(bits, 0), orc.branch_levels(bits, 1)


@pytest.mark.parametrize("bits", [1, 2])
def test_encoder_brute_force(bits):
    # Exhaustive search over every even-parity codeword of one block (b=1:
    # 2 x 128, b=2: 2 x 32768), cost = squared distance in the scaled
    # domain; the encoder's greedy + single-flip answer must be optimal and
    # match the brute-force argmin under the reference tie rules.
    rng = np.random.default_rng(bits)
    q = 1 << bits
    l0, l1 = _levels_for(bits)
    digits = np.array(list(itertools.product(range(q), repeat=8)), dtype=np.int64)
    digits = digits[digits.sum(axis=1) % 2 == 0]
    s = orc.lib().orc_fixed_scale(bits)
    P = 64
    u = rng.standard_normal((6, P))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u = u.astype(np.float32)
    dg, br = orc.encode(u, bits)
    rd = np.sqrt(P)
    for r in range(u.shape[0]):
        for blk in range(P // 8):
            v = u[r, blk * 8: blk * 8 + 8].astype(np.float64) * (s * rd)
            best = None
            for b_, L in ((0, l0), (1, l1)):
                cost = ((v[None, :] - L[digits]) ** 2).sum(axis=1)
                i = int(np.argmin(cost))
                if best is None or cost[i] < best[0] - 1e-12:
                    best = (cost[i], b_, digits[i])
            got_cost = ((v - (l1 if br[r, blk] else l0)[dg[r, blk * 8: blk * 8 + 8]]) ** 2).sum()
            assert got_cost <= best[0] + 1e-9, (r, blk)
            assert dg[r, blk * 8: blk * 8 + 8].sum() % 2 == 0


def test_encoder_parity_and_range():
    rng = np.random.default_rng(11)
    u = rng.standard_normal((40, 128))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    for b in range(1, 14):
        dg, br = orc.encode(u.astype(np.float32), b)
        assert dg.max() < (1 << b)
        assert np.all(dg.reshape(40, 16, 8).sum(axis=2) % 2 == 0)
        assert set(np.unique(br)) <= {0, 1}


# ---------------------------------------------------------------- coarse kernel
@pytest.mark.parametrize("blocks", cases.ACCUM_BLOCKS)
def test_coarse_accumulate_bit_exact(blocks):
    ref = gl("accum.npz")
    pay, lut = cases.accum_input(blocks)
    np.testing.assert_array_equal(orc.coarse_accumulate(pay[0], lut), ref[f"single_{blocks}"])
    np.testing.assert_array_equal(orc.coarse_accumulate(pay[0], lut), ref[f"pair_{blocks}"][0])
    np.testing.assert_array_equal(orc.coarse_accumulate(pay[1], lut), ref[f"pair_{blocks}"][1])
    # Closed form: lane t adds lut[half + lo] + lut[32 + half + hi].
    t = np.arange(32)
    half = (t >> 4) * 16
    want = np.zeros(32, np.int64)
    for b in range(blocks):
        h = pay[0][b]
        want += lut[b][half + (h & 15)].astype(np.int64) + lut[b][32 + half + (h >> 4)]
    np.testing.assert_array_equal(orc.coarse_accumulate(pay[0], lut), want)


# ---------------------------------------------------------------- whole index
def _case_index(name):
    return e8ax.read(os.path.join(G, f"{name}.e8ax"))


def test_manifest_inputs_unchanged():
    man = json.load(open(os.path.join(G, "manifest.json")))
    for name in cases.INDEX_CASES:
        data, cent, lab, _q = cases.mixture(name)
        assert cases.digest(data, cent, lab) == man["index"][name], name
    assert cases.digest(cases.encode_input()) == man["inputs"]["encode"]


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_build_bit_exact(name):
    n, dim, nlist, bits, seed, alpha, _e, _nq = cases.INDEX_CASES[name]
    data, cent, lab, _q = cases.mixture(name)
    ref = _case_index(name)
    got = orc.OracleIndex.build(data, cent, lab, bits, seed, alpha).arrays()
    for key in ("ids", "signs", "mean_rot", "centroid_rot", "centroid_n2", "cluster_begin", "arena_begin",
                "ref_arena_begin", "arena", "ref_arena"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=f"{name}:{key}")


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_search_bit_exact(name):
    ref = gl(f"{name}_search.npz")
    ix = orc.OracleIndex.from_e8ax(_case_index(name))
    queries = ref["queries"]
    for (k, nprobe, dis) in cases.search_grid(name):
        tag = f"k{k}_p{nprobe}_d{dis}"
        ids, d, st = ix.search(queries, k, nprobe, disable_prune=bool(dis))
        np.testing.assert_array_equal(ids, ref[f"{tag}_ids"], err_msg=tag)
        np.testing.assert_array_equal(d, ref[f"{tag}_dists"], err_msg=tag)
        assert int(st[0]) == int(ref[f"{tag}_stats"][:, 0].sum())
        assert int(st[1]) == int(ref[f"{tag}_stats"][:, 1].sum())
        for qi in range(queries.shape[0]):
            rq, lut, consts, probe, cd = ix.query_prep(queries[qi], k, nprobe)
            np.testing.assert_array_equal(rq, ref[f"{tag}_rotq"][qi])
            np.testing.assert_array_equal(lut, ref[f"{tag}_lut"][qi])
            np.testing.assert_array_equal(consts, ref[f"{tag}_consts"][qi])
            np.testing.assert_array_equal(probe, ref[f"{tag}_probe"][qi])
            np.testing.assert_array_equal(cd, ref[f"{tag}_pdist"][qi])


@pytest.mark.parametrize("name", ["d128b4", "d64b8", "d100b1"])
def test_search_equals_unpruned_estimator(name):
    # With prune disabled, search must be the estimator top-k over the
    # probed lists (reference e8q_test.cpp:281-373 asserts the same shape).
    ix = orc.OracleIndex.from_e8ax(_case_index(name))
    q = gl(f"{name}_search.npz")["queries"]
    nl = ix.nlist
    for nprobe in (2, nl):
        ids, d, _ = ix.search(q, 10, nprobe, disable_prune=True)
        ids2, d2 = ix.scan_all(q, 10, nprobe)
        np.testing.assert_array_equal(ids, ids2)
        np.testing.assert_array_equal(d, d2)


def test_estimator_against_float64():
    # The refined distance estimates ||q - x||^2: check the fixture index's
    # top-10 estimates against exact fp64 distances (a loose physical check
    # on g1/h3 and the decode -- 4-bit E8 error is a few percent).
    name = "d128b4"
    data, cent, lab, _q = cases.mixture(name)
    ref = gl(f"{name}_search.npz")
    q = ref["queries"]
    ids, d = ref["k10_p17_d1_ids"], ref["k10_p17_d1_dists"]
    for i in range(q.shape[0]):
        true = ((data[ids[i]].astype(np.float64) - q[i]) ** 2).sum(axis=1)
        rel = np.abs(d[i] - true) / true
        assert np.median(rel) < 0.05


@pytest.mark.parametrize("name", ["d128b4", "d100b1", "d128b2", "d64b9", "d768b4"])
def test_fp64_estimator_matches_fp32(name):
    # The fp64 restatement of the estimator (the MFMA form's error yardstick)
    # shares the pinned decode with the fp32 refine; the two differ only by
    # fp32 rounding of the products and sums: a wrong digit, branch, table
    # entry or sign would move a distance by far more than that.
    ix = orc.OracleIndex.from_e8ax(_case_index(name))
    q = gl(f"{name}_search.npz")["queries"]
    nl = ix.nlist
    for nprobe in (1, nl):
        ids, d = ix.scan_all(q, 10, nprobe)
        d64 = ix.dist64(q, nprobe, ids)
        ok = ids != np.iinfo(np.uint32).max
        assert np.all(np.isfinite(d64[ok])) and np.all(np.isnan(d64[~ok]))
        scale = np.abs(d64[ok]).max()
        err = np.abs(d[ok].astype(np.float64) - d64[ok])
        assert err.max() <= 1e-5 * scale, (name, err.max() / scale)
        # an id outside the probed lists has no distance
        if nprobe == 1 and ix.n > 0:
            far = np.setdiff1d(np.arange(ix.n, dtype=np.uint32), ids[0])[:1]
            probe = ix.query_prep(q[0], 10, 1)[3]
            d_far = ix.dist64(q[:1], 1, far[None, :])
            assert np.isnan(d_far[0, 0]) or far[0] in _members(ix, probe)


def _members(ix, lists):
    a = ix.arrays()
    out = []
    for c in lists:
        out.extend(a["ids"][a["cluster_begin"][c]:a["cluster_begin"][c + 1]].tolist())
    return set(out)
