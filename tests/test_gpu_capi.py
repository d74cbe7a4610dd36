"""Device library (C ABI) parity against the reference fixtures and the oracle.

Bit-exact: rotation, coarse integer sums, encoder digits/branches, every byte
of a device-built index (arena, banks, refine fields, g1/h3, ids), query LUT
bytes and constants, route distances and probe lists, and the plain search
ids and distances (the reference's pruned per-query schedule).
"""
import os

import numpy as np
import pytest

import capi
import cases
import e8ax
import oracle_py as orc
from tables import empty_arrays, layout, tables

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


def gl(name):
    return np.load(os.path.join(G, name))


@pytest.fixture(scope="module")
def ctx():
    c = capi.Ctx(0)
    yield c
    c.close()


def _bare_index(ctx, dim, bits, seed_signs):
    lay = layout(np.zeros(0, np.uint32), 1, bits, dim)
    arr = empty_arrays(0, dim, 1, bits, lay, seed_signs)
    return capi.DeviceIndex(ctx, arr, tables(bits, lay["padded"]))


@pytest.mark.parametrize("dim,seed", cases.ROTATE_CASES)
def test_rotate_bit_exact(ctx, dim, seed):
    ref = gl(f"rotate_{dim}.npz")
    ix = _bare_index(ctx, dim, 4, ref["signs"])
    print(f"rotate dim={dim}", flush=True)
    np.testing.assert_array_equal(ix.rotate(cases.rotate_input(dim)), ref["out"])


@pytest.mark.parametrize("blocks", cases.ACCUM_BLOCKS)
def test_coarse_accumulate_bit_exact(ctx, blocks):
    ref = gl("accum.npz")
    pay, lut = cases.accum_input(blocks)
    out = capi.coarse_accumulate(ctx, pay, lut)
    np.testing.assert_array_equal(out[0], ref[f"single_{blocks}"])
    np.testing.assert_array_equal(out, ref[f"pair_{blocks}"])


def test_coarse_accumulate_random(ctx):
    rng = np.random.default_rng(20260921)
    for blocks in (5, 129):
        pay = rng.integers(0, 256, size=(64, blocks, 32), dtype=np.uint8)
        lut = rng.integers(0, 256, size=(blocks, 64), dtype=np.uint8)
        out = capi.coarse_accumulate(ctx, pay, lut)
        for b in range(0, 64, 7):
            np.testing.assert_array_equal(out[b], orc.coarse_accumulate(pay[b], lut))


def test_encode_units_bit_exact(ctx):
    ref = gl("encode.npz")
    u = cases.encode_input()
    for b in range(1, 14):
        ix = _bare_index(ctx, cases.ENCODE_PADDED, b, orc.rotator_signs(cases.ENCODE_PADDED, 1))
        print(f"encode b={b}", flush=True)
        dg, br = ix.encode_units(u, orc.lib().orc_fixed_scale(b))
        np.testing.assert_array_equal(dg, ref[f"digits_b{b}"], err_msg=f"b={b}")
        np.testing.assert_array_equal(br, ref[f"branch_b{b}"], err_msg=f"b={b}")


def _fixture_index(ctx, name):
    a = e8ax.read(os.path.join(G, f"{name}.e8ax"))
    ix = capi.DeviceIndex(ctx, a, tables(a["bits"], a["padded"]))
    ix.upload()
    return a, ix


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_query_prep_bit_exact(ctx, name):
    a, ix = _fixture_index(ctx, name)
    ref = gl(f"{name}_search.npz")
    q = ref["queries"]
    for (k, nprobe, dis) in cases.search_grid(name)[:3]:
        tag = f"k{k}_p{nprobe}_d{dis}"
        print(f"query_prep {name} {tag}", flush=True)
        rq, lut, consts, probe, cd = ix.query_prep(q, nprobe)
        np.testing.assert_array_equal(rq, ref[f"{tag}_rotq"])
        np.testing.assert_array_equal(lut, ref[f"{tag}_lut"])
        np.testing.assert_array_equal(consts, ref[f"{tag}_consts"][:, :4])
        np.testing.assert_array_equal(probe, ref[f"{tag}_probe"])
        np.testing.assert_array_equal(cd, ref[f"{tag}_pdist"])


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_plain_search_bit_exact(ctx, name):
    a, ix = _fixture_index(ctx, name)
    ref = gl(f"{name}_search.npz")
    q = ref["queries"]
    for (k, nprobe, dis) in cases.search_grid(name):
        tag = f"k{k}_p{nprobe}_d{dis}"
        print(f"plain search {name} {tag}", flush=True)
        z = ref[f"{tag}_consts"][0, 4]
        ids, d, st = ix.search(q, k, nprobe, z, disable_prune=bool(dis))
        np.testing.assert_array_equal(ids, ref[f"{tag}_ids"], err_msg=tag)
        np.testing.assert_array_equal(d, ref[f"{tag}_dists"], err_msg=tag)
        assert st[0] == int(ref[f"{tag}_stats"][:, 0].sum())
        assert st[1] == int(ref[f"{tag}_stats"][:, 1].sum())


@pytest.mark.parametrize("name", list(cases.INDEX_CASES))
def test_device_build_bit_exact(ctx, name):
    n, dim, nlist, bits, seed, alpha, _e, _nq = cases.INDEX_CASES[name]
    data, cent, lab, _q = cases.mixture(name)
    ref = e8ax.read(os.path.join(G, f"{name}.e8ax"))
    lay = layout(lab, nlist, bits, dim)
    arr = empty_arrays(n, dim, nlist, bits, lay, orc.rotator_signs(dim, seed))
    ix = capi.DeviceIndex(ctx, arr, tables(bits, lay["padded"]))
    print(f"device build {name}", flush=True)
    ix.encode(data, cent, lab, lay["slot"])
    got = ix.download()
    for key in ("ids", "mean_rot", "centroid_rot", "centroid_n2", "arena", "ref_arena"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=f"{name}:{key}")


def test_random_vs_oracle(ctx):
    # A larger seeded case against the oracle: build bytes and pruned search.
    rng = np.random.default_rng(9)
    n, dim, nlist, bits = 20000, 128, 64, 4
    cent = (rng.standard_normal((nlist, dim)) * 2.0).astype(np.float32)
    lab = rng.integers(0, nlist, n).astype(np.uint32)
    data = (cent[lab] + rng.standard_normal((n, dim)).astype(np.float32)).astype(np.float32)
    q = (cent[rng.integers(0, nlist, 40)] + rng.standard_normal((40, dim)).astype(np.float32)).astype(np.float32)
    oi = orc.OracleIndex.build(data, cent, lab, bits, 42)
    oa = oi.arrays()
    lay = layout(lab, nlist, bits, dim)
    arr = empty_arrays(n, dim, nlist, bits, lay, orc.rotator_signs(dim, 42))
    ix = capi.DeviceIndex(ctx, arr, tables(bits, lay["padded"]))
    print("random build", flush=True)
    ix.encode(data, cent, lab, lay["slot"])
    got = ix.download()
    for key in ("ids", "mean_rot", "centroid_rot", "centroid_n2", "arena", "ref_arena"):
        np.testing.assert_array_equal(got[key], oa[key], err_msg=key)
    for k, nprobe, dis in ((100, 16, 0), (10, 8, 0), (10, 64, 1)):
        print(f"random search k={k} nprobe={nprobe} dis={dis}", flush=True)
        z = orc.lib().orc_ortho_quantile(1.0, max(4 * k, 32))
        ids, d, st = ix.search(q, k, nprobe, z, disable_prune=bool(dis))
        oids, od, ost = oi.search(q, k, nprobe, disable_prune=bool(dis))
        np.testing.assert_array_equal(ids, oids)
        np.testing.assert_array_equal(d, od)
        assert st == (int(ost[0]), int(ost[1]))
