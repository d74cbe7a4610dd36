"""ctypes binding of the device library's C ABI (include/e8q_gpu.h) for tests.

This is the binding a Python maintainer would add to call libe8q_gpu.so
directly; the product's own Python entry point is the pybind11 module
e8qlib (e8q-artifact_amd/e8qlib).
"""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "e8q-artifact_amd", "e8qlib", "libe8q_gpu.so")
HEADER = os.path.join(ROOT, "include", "e8q_gpu.h")
HEADERS = [HEADER, os.path.join(ROOT, "include", "me8q_gpu.h")]

vp = ctypes.c_void_p
u64 = ctypes.c_uint64


class IndexDesc(ctypes.Structure):
    _fields_ = [("n", u64), ("dim", u64), ("padded", u64), ("nlist", u64), ("bits", ctypes.c_int32),
                ("alpha", ctypes.c_float), ("scale", ctypes.c_double), ("signs", vp), ("refine_tab", vp),
                ("gate_mid", vp), ("levels", vp), ("cluster_begin", vp), ("arena_begin", vp),
                ("ref_arena_begin", vp)]


class IndexArrays(ctypes.Structure):
    _fields_ = [("ids", vp), ("mean_rot", vp), ("centroid_rot", vp), ("centroid_n2", vp), ("arena", vp),
                ("ref_arena", vp)]


class SearchStats(ctypes.Structure):
    _fields_ = [("candidates", u64), ("survivors", u64), ("batches", u64), ("scan_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("scanned", u64)]


class SearchArgs(ctypes.Structure):
    _fields_ = [("queries", vp), ("nq", u64), ("k", u64), ("nprobe", u64), ("nprobe_per_query", vp),
                ("z", ctypes.c_float), ("disable_prune", ctypes.c_int32), ("flush_every", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("queries_on_device", ctypes.c_int32), ("out_ids", vp),
                ("out_dists", vp), ("outputs_on_device", ctypes.c_int32)]


class CtxInfo(ctypes.Structure):
    _fields_ = [("qtile", ctypes.c_uint32), ("btile", ctypes.c_uint32), ("grid", ctypes.c_uint32),
                ("graphs", ctypes.c_uint32), ("tunings", u64), ("captures", u64), ("replays", u64)]


class M8CtxInfo(ctypes.Structure):
    _fields_ = [("mfma_grid", ctypes.c_uint32), ("refine_grid", ctypes.c_uint32), ("budget", ctypes.c_uint32),
                ("set_a", ctypes.c_uint32), ("tunings", u64), ("pool_per_query", u64)]


MODE_PLAIN, MODE_BATCHED = 0, 1

_lib = None


def declared_symbols():
    """Function names declared in include/e8q_gpu.h and include/me8q_gpu.h."""
    names = set()
    for h in HEADERS:
        text = open(h).read()
        names |= set(re.findall(r"^(?:e8q_status|uint64_t|const char\*)\s+(e8q_\w+)\(", text, re.M))
    return sorted(names)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libe8q_gpu.so not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.e8q_last_error.restype = ctypes.c_char_p
        u64_ret = ("e8q_index_device_bytes", "e8q_m8_device_bytes", "e8q_m8_row_bytes", "e8q_ctx_workspace_bytes",
                   "e8q_search_workspace_bytes", "e8q_m8_search_workspace_bytes")
        for name in declared_symbols():
            fn = getattr(L, name)
            if name in u64_ret:
                fn.restype = u64
            elif name != "e8q_last_error":
                fn.restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(f"e8q error {rc}: {lib().e8q_last_error().decode()}")


def p(a):
    return a.ctypes.data_as(vp) if a is not None else None


class Ctx:
    def __init__(self, device=0):
        self.h = vp()
        check(lib().e8q_ctx_create(device, ctypes.byref(self.h)))

    def set_graphs(self, on):
        check(lib().e8q_ctx_set_graphs(self.h, int(bool(on))))

    def set_batched_shape(self, qtile=0, btile=0, grid=0):
        check(lib().e8q_ctx_set_batched_shape(self.h, qtile, btile, grid))

    def info(self):
        i = CtxInfo()
        check(lib().e8q_ctx_get_info(self.h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in CtxInfo._fields_}

    def set_m8_grids(self, mfma=0, refine=0):
        check(lib().e8q_ctx_set_m8_grids(self.h, mfma, refine))

    def m8_info(self):
        i = M8CtxInfo()
        check(lib().e8q_ctx_get_m8_info(self.h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in M8CtxInfo._fields_}

    def reserve(self, nbytes):
        check(lib().e8q_ctx_reserve(self.h, ctypes.c_uint64(int(nbytes))))

    def share_tuning(self, other):
        check(lib().e8q_ctx_share_tuning(self.h, other.h))

    def workspace(self):
        return int(lib().e8q_ctx_workspace_bytes(self.h))

    def close(self):
        if self.h:
            lib().e8q_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DeviceIndex:
    """Device index from E8AX arrays plus the static tables (levels etc.)."""

    def __init__(self, ctx, arrays, tables):
        self.ctx = ctx
        self.a = arrays
        self.t = tables
        self._keep = []
        d = IndexDesc()
        d.n, d.dim, d.padded, d.nlist, d.bits = (arrays["n"], arrays["dim"], arrays["padded"], arrays["nlist"],
                                                 arrays["bits"])
        d.alpha = arrays.get("alpha", 1.0)
        d.scale = tables["scale"]
        for f in ("signs", "refine_tab", "gate_mid", "levels"):
            arr = np.ascontiguousarray(tables[f] if f != "signs" else arrays["signs"])
            self._keep.append(arr)
            setattr(d, f, p(arr))
        for f in ("cluster_begin", "arena_begin", "ref_arena_begin"):
            arr = np.ascontiguousarray(arrays[f], np.uint64)
            self._keep.append(arr)
            setattr(d, f, p(arr))
        self.h = vp()
        check(lib().e8q_index_create(ctx.h, ctypes.byref(d), ctypes.byref(self.h)))

    def upload(self):
        h = IndexArrays()
        keep = []
        for f in ("ids", "mean_rot", "centroid_rot", "centroid_n2", "arena", "ref_arena"):
            arr = np.ascontiguousarray(self.a[f])
            keep.append(arr)
            setattr(h, f, p(arr))
        check(lib().e8q_index_upload(self.ctx.h, self.h, ctypes.byref(h)))

    def encode(self, base, cent, cid, slot):
        base = np.ascontiguousarray(base, np.float32)
        cent = np.ascontiguousarray(cent, np.float32)
        cid = np.ascontiguousarray(cid, np.uint32)
        slot = np.ascontiguousarray(slot, np.uint32)
        check(lib().e8q_index_encode(self.ctx.h, self.h, p(base), p(cent), p(cid), p(slot), 0))

    def download(self):
        a = self.a
        out = {k: np.empty_like(np.asarray(a[k])) for k in ("ids", "mean_rot", "centroid_rot", "centroid_n2",
                                                            "arena", "ref_arena")}
        h = IndexArrays()
        for f, arr in out.items():
            setattr(h, f, p(arr))
        check(lib().e8q_index_download(self.ctx.h, self.h, ctypes.byref(h)))
        return out

    def search(self, queries, k, nprobe, z, disable_prune=False, mode=MODE_PLAIN, nprobe_per_query=None,
               flush_every=32):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        dists = np.empty((nq, k), np.float32)
        npq = None if nprobe_per_query is None else np.ascontiguousarray(nprobe_per_query, np.uint32)
        s = SearchArgs()
        s.queries, s.nq, s.k, s.nprobe = p(q), nq, k, nprobe
        s.nprobe_per_query = p(npq)
        s.z, s.disable_prune, s.flush_every, s.mode = z, int(bool(disable_prune)), flush_every, mode
        s.queries_on_device, s.outputs_on_device = 0, 0
        s.out_ids, s.out_dists = p(ids), p(dists)
        st = SearchStats()
        check(lib().e8q_search(self.ctx.h, self.h, ctypes.byref(s), ctypes.byref(st)))
        return ids, dists, (st.candidates, st.survivors)

    def search_full(self, queries, k, nprobe, z, disable_prune=False, nprobe_per_query=None, flush_every=32,
                    static=False, device_ptrs=None):
        """A plain search through e8q_search or (static=True) e8q_search_static;
        returns ids, dists and the whole stats struct as a dict.  device_ptrs =
        (queries, out_ids, out_dists) device addresses: queries and outputs
        stay on the device (ids / dists are then None)."""
        s = SearchArgs()
        npq = None if nprobe_per_query is None else np.ascontiguousarray(nprobe_per_query, np.uint32)
        if device_ptrs is None:
            q = np.ascontiguousarray(queries, np.float32)
            nq = q.shape[0]
            ids = np.empty((nq, k), np.uint32)
            dists = np.empty((nq, k), np.float32)
            s.queries, s.out_ids, s.out_dists = p(q), p(ids), p(dists)
            s.queries_on_device, s.outputs_on_device = 0, 0
        else:
            nq = int(queries)
            ids = dists = None
            s.queries, s.out_ids, s.out_dists = device_ptrs
            s.queries_on_device, s.outputs_on_device = 1, 1
        s.nq, s.k, s.nprobe = nq, k, nprobe
        s.nprobe_per_query = p(npq)
        s.z, s.disable_prune, s.flush_every, s.mode = z, int(bool(disable_prune)), flush_every, MODE_PLAIN
        st = SearchStats()
        fn = lib().e8q_search_static if static else lib().e8q_search
        check(fn(self.ctx.h, self.h, ctypes.byref(s), ctypes.byref(st)))
        return ids, dists, {f: getattr(st, f) for f, _ in SearchStats._fields_}

    def workspace_bytes(self, nq, k, max_nprobe):
        return int(lib().e8q_search_workspace_bytes(self.h, nq, k, max_nprobe))

    def estimate(self, queries, qidx, ids):
        q = np.ascontiguousarray(queries, np.float32)
        qi, ii = np.ascontiguousarray(qidx, np.uint32), np.ascontiguousarray(ids, np.uint32)
        out = np.empty(ii.size, np.float32)
        check(lib().e8q_estimate(self.ctx.h, self.h, p(q), u64(q.shape[0]), p(qi), p(ii), u64(ii.size), p(out)))
        return out

    def probe_bound_w(self, queries, base_rows, qidx, ids):
        q = np.ascontiguousarray(queries, np.float32)
        b = np.ascontiguousarray(base_rows, np.float32)
        qi, ii = np.ascontiguousarray(qidx, np.uint32), np.ascontiguousarray(ids, np.uint32)
        out = np.empty(ii.size, np.float64)
        check(lib().e8q_probe_bound_w(self.ctx.h, self.h, p(q), u64(q.shape[0]), p(b), p(qi), p(ii), u64(ii.size),
                                      p(out)))
        return out

    def exhaustive_topk(self, queries, k):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        check(lib().e8q_exhaustive_topk(self.ctx.h, self.h, p(q), u64(nq), u64(k), p(ids), p(d)))
        return ids, d

    def query_prep(self, queries, nprobe):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        P, L = self.a["padded"], self.a["nlist"]
        rq = np.empty((nq, P), np.float32)
        lut = np.empty((nq, P // 8 * 64), np.uint8)
        consts = np.empty((nq, 4), np.float32)
        take = min(nprobe, L)
        probe = np.empty((nq, take), np.uint32)
        cd = np.empty((nq, L), np.float32)
        check(lib().e8q_debug_query_prep(self.ctx.h, self.h, p(q), nq, nprobe, p(rq), p(lut), p(consts), p(probe),
                                         p(cd)))
        return rq, lut, consts, probe, cd

    def rotate(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty((x.shape[0], self.a["padded"]), np.float32)
        check(lib().e8q_debug_rotate(self.ctx.h, self.h, p(x), x.shape[0], p(out)))
        return out

    def encode_units(self, unit, scale):
        unit = np.ascontiguousarray(unit, np.float32)
        rows, P = unit.shape
        dg = np.empty((rows, P), np.uint16)
        br = np.empty((rows, P // 8), np.uint8)
        check(lib().e8q_debug_encode(self.ctx.h, self.h, p(unit), rows, ctypes.c_double(scale), p(dg), p(br)))
        return dg, br

    def close(self):
        if self.h:
            lib().e8q_index_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def coarse_accumulate(ctx, rowa, lut):
    rowa = np.ascontiguousarray(rowa, np.uint8)    # [nbatch, blocks, 32]
    lut = np.ascontiguousarray(lut, np.uint8)      # [blocks, 64]
    nb = rowa.shape[0]
    out = np.empty((nb, 32), np.int32)
    check(lib().e8q_debug_coarse_accumulate(ctx.h, p(rowa), nb, p(lut), lut.shape[0], p(out)))
    return out
