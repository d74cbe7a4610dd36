"""ctypes binding of the ME8Q half of the device library's C ABI
(include/me8q_gpu.h) for tests: the binding a Python maintainer would add to
call libe8q_gpu.so directly (the product's Python entry is e8qlib.Me8qIndex)."""
import ctypes
import os

import numpy as np

import capi
from capi import check, lib, p, u64, vp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class M8Desc(ctypes.Structure):
    _fields_ = [("n", u64), ("dim", u64), ("padded", u64), ("nlist", u64), ("nseg", ctypes.c_uint32),
                ("seg_len", vp), ("seg_bits", vp), ("signs", vp), ("levels", vp), ("fixed_scale", vp),
                ("list_begin", vp), ("drop_energy", ctypes.c_int32)]


class M8Arrays(ctypes.Structure):
    _fields_ = [("ids", vp), ("centroids", vp), ("codes", vp)]


class M8SearchArgs(ctypes.Structure):
    _fields_ = [("queries", vp), ("nq", u64), ("k", u64), ("nprobe", u64), ("nprobe_per_query", vp),
                ("z", ctypes.c_float), ("queries_on_device", ctypes.c_int32), ("out_ids", vp), ("out_dists", vp),
                ("outputs_on_device", ctypes.c_int32), ("mode", ctypes.c_int32), ("batched_budget", ctypes.c_uint32),
                ("batched_set_a", ctypes.c_uint32)]


class M8Stats(ctypes.Structure):
    _fields_ = [("candidates", u64), ("survivors", u64), ("bits_scored", u64), ("bits_total", u64),
                ("pruned", u64), ("scan_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("batches", u64),
                ("refined", u64), ("groups", u64), ("form", ctypes.c_int32), ("overflowed", u64),
                ("pairs_scored", u64), ("items", u64), ("refine_ms", ctypes.c_double * 2),
                ("refine_launches", u64 * 2), ("refine_bytes", u64 * 2)]

PLAIN, BATCHED = 0, 1


def tables():
    """Level tables (b = 1..13, branch 0 then 1) and FixedScale, as written
    by the reference library."""
    return (np.ascontiguousarray(np.load(os.path.join(GOLDEN, "levels.npy")), np.float64),
            np.ascontiguousarray(np.load(os.path.join(GOLDEN, "fixed_scale.npy")), np.float64))


class M8Index:
    def __init__(self, ctx, n, dim, nlist, plan, signs, list_begin, drop_energy):
        self.ctx = ctx
        self.n, self.dim, self.nlist = n, dim, nlist
        self.padded = (dim + 63) // 64 * 64
        self.plan = list(plan)
        levels, fixed = tables()
        seg_len = np.array([s[0] for s in plan], np.uint64)
        seg_bits = np.array([s[1] for s in plan], np.int32)
        sg = np.concatenate([np.asarray(s, np.uint8) for s in signs]) if signs else np.zeros(1, np.uint8)
        lb = np.ascontiguousarray(list_begin, np.uint64)
        self._keep = [levels, fixed, seg_len, seg_bits, sg, lb]
        d = M8Desc()
        d.n, d.dim, d.padded, d.nlist, d.nseg = n, dim, self.padded, nlist, len(plan)
        d.seg_len, d.seg_bits, d.signs, d.levels, d.fixed_scale, d.list_begin = (p(seg_len), p(seg_bits), p(sg),
                                                                                 p(levels), p(fixed), p(lb))
        d.drop_energy = int(bool(drop_energy))
        self.h = vp()
        check(lib().e8q_m8_create(ctx.h, ctypes.byref(d), ctypes.byref(self.h)))
        self.row_bytes = int(lib().e8q_m8_row_bytes(self.h))
        self.list_begin = lb
        self.plane_cols = sum(s[0] // 8 for s in plan if s[1] > 0)
        self.nrun = sum(1 for s in plan if s[1] > 0)

    @classmethod
    def from_file(cls, ctx, f):
        ix = cls(ctx, f["n"], f["dim"], f["nlist"], f["plan"], f["signs"], f["list_begin"], f["drop_energy"])
        ix.upload(f["ids"], f["centroids"], f["codes"])
        return ix

    def upload(self, ids, centroids, codes):
        a = M8Arrays()
        keep = [np.ascontiguousarray(ids, np.uint32), np.ascontiguousarray(centroids, np.float32),
                np.ascontiguousarray(codes, np.uint8)]
        a.ids, a.centroids, a.codes = (p(x) for x in keep)
        check(lib().e8q_m8_upload(self.ctx.h, self.h, ctypes.byref(a)))

    def encode(self, base, centroids_padded, cid, rank):
        keep = [np.ascontiguousarray(base, np.float32), np.ascontiguousarray(centroids_padded, np.float32),
                np.ascontiguousarray(cid, np.uint32), np.ascontiguousarray(rank, np.uint32)]
        check(lib().e8q_m8_encode(self.ctx.h, self.h, *(p(x) for x in keep), 0))

    def download(self):
        ids = np.empty(self.n, np.uint32)
        cent = np.empty((self.nlist, self.padded), np.float32)
        codes = np.empty(self.n * self.row_bytes, np.uint8)
        a = M8Arrays()
        a.ids, a.centroids, a.codes = p(ids), p(cent), p(codes)
        check(lib().e8q_m8_download(self.ctx.h, self.h, ctypes.byref(a)))
        return ids, cent, codes

    def fast(self, c):
        sz = np.zeros(3, np.uint64)
        check(lib().e8q_m8_debug_fast(self.ctx.h, self.h, c, p(sz), None, None, None, None, None))
        plane = np.empty(int(sz[0]), np.uint8)
        refine = np.empty(int(sz[1]), np.uint8)
        gain = np.empty(int(sz[2]), np.float32)
        vpc = np.empty(int(sz[2]), np.float32)
        rcn2 = np.empty(max(self.nrun, 1), np.float64)
        check(lib().e8q_m8_debug_fast(self.ctx.h, self.h, c, p(sz), p(plane), p(refine), p(gain), p(vpc), p(rcn2)))
        return plane, refine, gain, vpc, rcn2[:self.nrun]

    def centroids_rot(self):
        crot = np.empty((self.nlist, self.padded), np.float32)
        cn2 = np.empty(self.nlist, np.float32)
        check(lib().e8q_m8_debug_centroids(self.ctx.h, self.h, p(crot), p(cn2)))
        return crot, cn2

    def query(self, queries, nprobe):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        take = min(nprobe, self.nlist)
        rq = np.empty((nq, self.padded), np.float32)
        lut = np.empty((nq, self.plane_cols * 64), np.uint8)
        md = np.empty((nq, 2), np.float32)
        probe = np.empty((nq, take), np.uint32)
        pd = np.empty((nq, take), np.float32)
        check(lib().e8q_m8_debug_query(self.ctx.h, self.h, p(q), nq, nprobe, p(rq), p(lut), p(md), p(probe), p(pd)))
        return rq, lut, md, probe, pd

    def _args(self, q, k, nprobe, z, npq, mode, budget, ids, d, set_a=0):
        s = M8SearchArgs()
        s.queries, s.nq, s.k, s.nprobe, s.nprobe_per_query = p(q), q.shape[0], k, nprobe, p(npq)
        s.z, s.queries_on_device, s.outputs_on_device = z, 0, 0
        s.out_ids, s.out_dists = p(ids), p(d)
        s.mode, s.batched_budget, s.batched_set_a = mode, budget, set_a
        return s

    def workspace_bytes(self, nq, k, max_nprobe, mode=PLAIN):
        return int(lib().e8q_m8_search_workspace_bytes(self.h, u64(nq), u64(k), u64(max_nprobe), mode))

    def search(self, queries, k, nprobe, z, nprobe_per_query=None, mode=PLAIN, budget=0, set_a=0):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        st = np.empty((nq, 5), np.uint64)
        npq = None if nprobe_per_query is None else np.ascontiguousarray(nprobe_per_query, np.uint32)
        s = self._args(q, k, nprobe, z, npq, mode, budget, ids, d, set_a)
        tot = M8Stats()
        check(lib().e8q_m8_search(self.ctx.h, self.h, ctypes.byref(s), ctypes.byref(tot), p(st)))
        return ids, d, st, tot

    def estimate(self, queries, qidx, ids):
        q = np.ascontiguousarray(queries, np.float32)
        qi, ii = np.ascontiguousarray(qidx, np.uint32), np.ascontiguousarray(ids, np.uint32)
        out = np.empty(ii.size, np.float32)
        check(lib().e8q_m8_estimate(self.ctx.h, self.h, p(q), u64(q.shape[0]), p(qi), p(ii), u64(ii.size), p(out)))
        return out

    def exhaustive_topk(self, queries, k):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        check(lib().e8q_m8_exhaustive_topk(self.ctx.h, self.h, p(q), u64(nq), u64(k), p(ids), p(d)))
        return ids, d

    def search_ivf(self, queries, k, nprobe):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        st = np.empty((nq, 6), np.uint64)
        check(lib().e8q_m8_search_ivf(self.ctx.h, self.h, p(q), u64(nq), u64(k), u64(nprobe), p(ids), p(d), p(st)))
        return ids, d, st

    def batch_checksum(self, queries, k, nprobe, z, nprobe_per_query=None, budget=0):
        """Per-query checksum of the batched form's integer coarse sums."""
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        npq = None if nprobe_per_query is None else np.ascontiguousarray(nprobe_per_query, np.uint32)
        s = self._args(q, k, nprobe, z, npq, BATCHED, budget, ids, d)
        chk = np.empty(nq, np.uint32)
        check(lib().e8q_m8_debug_batch_checksum(self.ctx.h, self.h, ctypes.byref(s), p(chk)))
        return chk

    def close(self):
        if self.h:
            lib().e8q_m8_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def ranks(cid, nlist):
    """Position of each row inside its list (rows in file order)."""
    cid = np.asarray(cid, np.int64)
    rank = np.empty(cid.size, np.uint32)
    seen = np.zeros(nlist, np.int64)
    for i, c in enumerate(cid):
        rank[i] = seen[c]
        seen[c] += 1
    return rank


def list_begin(cid, nlist):
    cnt = np.bincount(np.asarray(cid, np.int64), minlength=nlist)
    return np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint64)
