"""ctypes front of the oracle (oracle/e8q_oracle.cpp) -- test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liborc_e8q.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i32, u64, f32, f64 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.c_float, ctypes.c_double)
        L.orc_fixed_scale.restype = f64
        L.orc_fixed_scale.argtypes = [i32]
        L.orc_branch_levels.argtypes = [i32, i32, vp]
        L.orc_scan_class_means.argtypes = [i32, vp]
        L.orc_refine_table.argtypes = [i32, sz, vp]
        L.orc_rotator_signs.restype = sz
        L.orc_rotator_signs.argtypes = [sz, u64, vp]
        L.orc_rotate.argtypes = [sz, u64, vp, sz, vp]
        L.orc_encode.argtypes = [vp, sz, i32, i32, f64, vp, vp]
        L.orc_coarse_accumulate.argtypes = [vp, vp, sz, vp]
        L.orc_ortho_quantile.restype = f32
        L.orc_ortho_quantile.argtypes = [f64, sz]
        L.orc_build.restype = vp
        L.orc_build.argtypes = [vp, vp, vp, sz, sz, sz, i32, u64, f32]
        L.orc_free.argtypes = [vp]
        L.orc_index_info.argtypes = [vp, vp]
        L.orc_index_arrays.argtypes = [vp] + [vp] * 10
        L.orc_from_arrays.restype = vp
        L.orc_from_arrays.argtypes = [u64, u64, u64, i32] + [vp] * 10 + [f32]
        L.orc_set_alpha.argtypes = [vp, f32]
        L.orc_search.restype = i32
        L.orc_search.argtypes = [vp, vp, sz, sz, sz, i32, i32, vp, vp, vp]
        L.orc_query_prep.argtypes = [vp, vp, sz, sz, vp, vp, vp, vp, vp]
        L.orc_scan_all.argtypes = [vp, vp, sz, sz, sz, vp, vp]
        L.orc_dist64.argtypes = [vp, vp, sz, sz, vp, sz, vp]
        L.orc_estimate.restype = i32
        L.orc_estimate.argtypes = [vp, vp, vp, vp, sz, vp]
        L.orc_probe_bound_w.restype = i32
        L.orc_probe_bound_w.argtypes = [vp, vp, vp, vp, vp, sz, vp]
        L.orc_exhaustive_topk.argtypes = [vp, vp, sz, sz, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def c32(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


def branch_levels(bits, branch):
    out = np.empty(1 << bits, np.float64)
    lib().orc_branch_levels(bits, branch, _p(out))
    return out


def class_means(bits):
    out = np.empty(2, np.float64)
    lib().orc_scan_class_means(bits, _p(out))
    return out


def rotator_signs(dim, seed):
    n = lib().orc_rotator_signs(dim, seed, None)
    out = np.empty(n, np.uint8)
    lib().orc_rotator_signs(dim, seed, _p(out))
    return out


def rotate(dim, seed, x):
    x = c32(x)
    P = (dim + 63) // 64 * 64
    out = np.empty((x.shape[0], P), np.float32)
    lib().orc_rotate(dim, seed, _p(x), x.shape[0], _p(out))
    return out


def encode(unit, bits, scale=None):
    unit = c32(unit)
    rows, P = unit.shape
    s = lib().orc_fixed_scale(bits) if scale is None else scale
    dg = np.empty((rows, P), np.uint16)
    br = np.empty((rows, P // 8), np.uint8)
    lib().orc_encode(_p(unit), rows, P, bits, s, _p(dg), _p(br))
    return dg, br


def coarse_accumulate(rowa, lut):
    rowa = np.ascontiguousarray(rowa, np.uint8)
    lut = np.ascontiguousarray(lut, np.uint8)
    out = np.empty(32, np.int32)
    lib().orc_coarse_accumulate(_p(rowa), _p(lut), lut.shape[0], _p(out))
    return out


class OracleIndex:
    def __init__(self, handle):
        self.h = handle
        info = np.empty(10, np.uint64)
        lib().orc_index_info(self.h, _p(info))
        (self.n, self.dim, self.padded, self.nlist, self.bits, self.blocks, self.arena_bytes,
         self.ref_bytes, self.batch_bytes, self.ref_batch_bytes) = [int(v) for v in info]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    @classmethod
    def build(cls, data, cent, ids, bits, seed, alpha=0.0):
        data, cent = c32(data), c32(cent)
        ids = np.ascontiguousarray(ids, np.uint32)
        h = lib().orc_build(_p(data), _p(cent), _p(ids), data.shape[0], data.shape[1], cent.shape[0], bits,
                            seed, alpha)
        if not h:
            raise ValueError("oracle build failed")
        return cls(h)

    @classmethod
    def from_e8ax(cls, ix):
        h = lib().orc_from_arrays(ix["n"], ix["dim"], ix["nlist"], ix["bits"], _p(ix["ids"]), _p(ix["signs"]),
                                  _p(ix["mean_rot"]), _p(ix["centroid_rot"]), _p(ix["centroid_n2"]),
                                  _p(ix["cluster_begin"]), _p(ix["arena_begin"]), _p(ix["arena"]),
                                  _p(ix["ref_arena_begin"]), _p(ix["ref_arena"]), ix["alpha"])
        return cls(h)

    def arrays(self):
        P, L = self.padded, self.nlist
        a = dict(n=self.n, dim=self.dim, nlist=L, bits=self.bits, padded=P)
        a["ids"] = np.empty(self.n, np.uint32)
        a["signs"] = np.empty(4 * P // 8, np.uint8)
        a["mean_rot"] = np.empty(P, np.float32)
        a["centroid_rot"] = np.empty((L, P), np.float32)
        a["centroid_n2"] = np.empty(L, np.float32)
        a["cluster_begin"] = np.empty(L + 1, np.uint64)
        a["arena_begin"] = np.empty(L + 1, np.uint64)
        a["arena"] = np.empty(self.arena_bytes, np.uint8)
        a["ref_arena_begin"] = np.empty(L + 1, np.uint64)
        a["ref_arena"] = np.empty(self.ref_bytes, np.uint8)
        lib().orc_index_arrays(self.h, *[_p(a[k]) for k in ("ids", "signs", "mean_rot", "centroid_rot",
                                                           "centroid_n2", "cluster_begin", "arena_begin",
                                                           "arena", "ref_arena_begin", "ref_arena")])
        return a

    def set_alpha(self, alpha):
        lib().orc_set_alpha(self.h, alpha)

    def search(self, queries, k, nprobe, disable_prune=False, threads=1):
        q = c32(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        st = np.zeros(2, np.uint64)
        rc = lib().orc_search(self.h, _p(q), nq, k, nprobe, int(bool(disable_prune)), threads, _p(ids), _p(d),
                              _p(st))
        if rc != 0:
            raise ValueError("bad search arguments")
        return ids, d, st

    def query_prep(self, query, k, nprobe):
        q = c32(query)
        rq = np.empty(self.padded, np.float32)
        lut = np.empty(self.blocks * 64, np.uint8)
        consts = np.empty(5, np.float32)
        take = min(nprobe, self.nlist)
        probe = np.empty(take, np.uint32)
        cd = np.empty(self.nlist, np.float32)
        lib().orc_query_prep(self.h, _p(q), k, nprobe, _p(rq), _p(lut), _p(consts), _p(probe), _p(cd))
        return rq, lut, consts, probe, cd

    def scan_all(self, queries, k, nprobe):
        q = c32(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        lib().orc_scan_all(self.h, _p(q), nq, k, nprobe, _p(ids), _p(d))
        return ids, d

    def estimate(self, queries, qidx, ids):
        """Estimate of vector ids[i] for queries[qidx[i]] (KeyError: unknown id)."""
        q = c32(queries)
        qi, ii = c32(qidx, np.uint32), c32(ids, np.uint32)
        out = np.empty(ii.size, np.float32)
        if lib().orc_estimate(self.h, _p(q), _p(qi), _p(ii), ii.size, _p(out)) != 0:
            raise KeyError("unknown id")
        return out

    def probe_bound_w(self, queries, base_rows, qidx, ids):
        q, b = c32(queries), c32(base_rows)
        qi, ii = c32(qidx, np.uint32), c32(ids, np.uint32)
        out = np.empty(ii.size, np.float64)
        if lib().orc_probe_bound_w(self.h, _p(q), _p(b), _p(qi), _p(ii), ii.size, _p(out)) != 0:
            raise KeyError("unknown id")
        return out

    def exhaustive_topk(self, queries, k):
        q = c32(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        lib().orc_exhaustive_topk(self.h, _p(q), nq, k, _p(ids), _p(d))
        return ids, d

    def dist64(self, queries, nprobe, ids):
        """fp64 estimator distances of the given ids (NaN: not in the probed lists)."""
        q = c32(queries)
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.empty(ids.shape, np.float64)
        lib().orc_dist64(self.h, _p(q), q.shape[0], nprobe, _p(ids), ids.shape[1], _p(out))
        return out


# ---------------------------------------------------------------------------
# ME8Q (oracle/me8q_oracle.cpp)

def _m8_lib():
    L = lib()
    if not getattr(L, "_m8_ready", False):
        vp, sz, i32, u64, f32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64, ctypes.c_float
        L.orc_m8_build.restype = vp
        L.orc_m8_build.argtypes = [vp, vp, vp, sz, sz, sz, f32, u64, f32, ctypes.c_char_p, vp, i32]
        L.orc_m8_free.argtypes = [vp]
        L.orc_m8_info.argtypes = [vp, vp]
        L.orc_m8_image.argtypes = [vp, vp]
        L.orc_m8_plan.argtypes = [vp, vp]
        L.orc_m8_fast_sizes.argtypes = [vp, sz, vp]
        L.orc_m8_fast.argtypes = [vp, sz, vp, vp, vp, vp, vp]
        L.orc_m8_centroids.argtypes = [vp, vp, vp]
        L.orc_m8_query.argtypes = [vp, vp, sz, vp, vp, vp, vp, vp]
        L.orc_m8_search.restype = i32
        L.orc_m8_search.argtypes = [vp, vp, sz, sz, sz, i32, vp, vp, vp]
        L.orc_m8_search_z.restype = i32
        L.orc_m8_search_z.argtypes = [vp, vp, sz, sz, sz, i32, f32, vp, vp, vp]
        L.orc_m8_estimate.restype = i32
        L.orc_m8_estimate.argtypes = [vp, vp, vp, vp, sz, vp]
        L.orc_m8_exhaustive_topk.argtypes = [vp, vp, sz, sz, vp, vp]
        L.orc_m8_search_ivf.argtypes = [vp, vp, sz, sz, sz, vp, vp, vp]
        L._m8_ready = True
    return L


class OracleMe8q:
    """Oracle ME8Q index (build + FastScan search), reference semantics."""

    def __init__(self, h):
        self.h = h
        info = np.zeros(11, np.uint64)
        _m8_lib().orc_m8_info(h, _p(info))
        (self.n, self.dim, self.padded, self.nlist, self.nrun, self.plane_cols, self.refine_stride, self.stride,
         self.route_dim, self.image_bytes, plan_len) = (int(v) for v in info)
        plan = np.zeros((plan_len, 2), np.int64)
        _m8_lib().orc_m8_plan(h, _p(plan))
        self.plan = [(int(a), int(b)) for a, b in plan]

    @classmethod
    def build(cls, data, cent, lab, avg_bits=4.0, seed=0, prune_alpha=0.0, plan_spec="", variance=None,
              drop_energy=True):
        data, cent, lab = c32(data), c32(cent), c32(lab, np.uint32)
        var = c32(variance) if variance is not None else None
        h = _m8_lib().orc_m8_build(_p(data), _p(cent), _p(lab), data.shape[0], data.shape[1], cent.shape[0],
                                   float(avg_bits), int(seed), float(prune_alpha), plan_spec.encode(), _p(var),
                                   int(drop_energy))
        if not h:
            raise ValueError("oracle ME8Q build failed")
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None):
            _m8_lib().orc_m8_free(self.h)
            self.h = None

    def image(self):
        out = np.empty(self.image_bytes, np.uint8)
        _m8_lib().orc_m8_image(self.h, _p(out))
        return out

    def fast(self, c):
        sz = np.zeros(3, np.uint64)
        _m8_lib().orc_m8_fast_sizes(self.h, c, _p(sz))
        plane = np.empty(int(sz[0]), np.uint8)
        refine = np.empty(int(sz[1]), np.uint8)
        gain = np.empty(int(sz[2]), np.float32)
        vpc = np.empty(int(sz[2]), np.float32)
        rcn2 = np.empty(self.nrun, np.float64)
        _m8_lib().orc_m8_fast(self.h, c, _p(plane), _p(refine), _p(gain), _p(vpc), _p(rcn2))
        return plane, refine, gain, vpc, rcn2

    def centroids(self):
        crot = np.empty((self.nlist, self.padded), np.float32)
        cn2 = np.empty(self.nlist, np.float32)
        _m8_lib().orc_m8_centroids(self.h, _p(crot), _p(cn2))
        return crot, cn2

    def query(self, q, nprobe):
        q = c32(q)
        take = min(nprobe, self.nlist)
        rq = np.empty(self.padded, np.float32)
        lut = np.empty(self.plane_cols * 64, np.uint8)
        md = np.empty(2, np.float32)
        probe = np.empty(take, np.uint32)
        pd = np.empty(take, np.float32)
        _m8_lib().orc_m8_query(self.h, _p(q), nprobe, _p(rq), _p(lut), _p(md), _p(probe), _p(pd))
        return rq, lut, md, probe, pd

    def search(self, queries, k, nprobe, threads=8, z=None):
        """The reference's search; z overrides the union-bound quantile (the
        index's own when None; +inf prunes nothing, giving the exact
        estimator top-k over the probed lists)."""
        q = c32(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        st = np.empty((nq, 5), np.uint64)
        if z is None:
            _m8_lib().orc_m8_search(self.h, _p(q), nq, k, nprobe, threads, _p(ids), _p(d), _p(st))
        else:
            _m8_lib().orc_m8_search_z(self.h, _p(q), nq, k, nprobe, threads, float(z), _p(ids), _p(d), _p(st))
        return ids, d, st

    def estimate(self, queries, qidx, ids):
        """Estimate of vector ids[i] for queries[qidx[i]] (KeyError: unknown id)."""
        q = c32(queries)
        qi, ii = c32(qidx, np.uint32), c32(ids, np.uint32)
        out = np.empty(ii.size, np.float32)
        if _m8_lib().orc_m8_estimate(self.h, _p(q), _p(qi), _p(ii), ii.size, _p(out)) != 0:
            raise KeyError("unknown id")
        return out

    def exhaustive_topk(self, queries, k):
        q = c32(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        _m8_lib().orc_m8_exhaustive_topk(self.h, _p(q), nq, k, _p(ids), _p(d))
        return ids, d

    def search_ivf(self, queries, k, nprobe):
        """SearchIVF: ids, distances and per-query (candidates, clusters,
        bits_scored, bits_total, survivors, pruned)."""
        q = c32(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        st = np.empty((nq, 6), np.uint64)
        _m8_lib().orc_m8_search_ivf(self.h, _p(q), nq, k, nprobe, _p(ids), _p(d), _p(st))
        return ids, d, st
