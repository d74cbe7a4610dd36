"""Reader/writer for the E8AX index file (reference e8q_io.cpp:35-211)."""
import struct

import numpy as np

MAGIC = 0x58413845
OPT_MAGIC = 0x514F5045


def padded_dim(dim):
    return (dim + 63) // 64 * 64


def read(path):
    with open(path, "rb") as f:
        buf = f.read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from(fmt, buf, off)
        off += struct.calcsize(fmt)
        return v

    def arr(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off).copy()
        off += a.nbytes
        return a

    (magic,) = take("<I")
    assert magic == MAGIC, "not an E8AX file"
    n, dim, nlist = take("<QQQ")
    (bits,) = take("<i")
    (rot_bytes,) = take("<Q")
    P = padded_dim(dim)
    ix = dict(n=n, dim=dim, nlist=nlist, bits=bits, padded=P)
    ix["ids"] = arr("<u4", n)
    ix["signs"] = arr("u1", rot_bytes)
    ix["mean_rot"] = arr("<f4", P)
    ix["centroid_rot"] = arr("<f4", nlist * P).reshape(nlist, P)
    ix["centroid_n2"] = arr("<f4", nlist)
    ix["cluster_begin"] = arr("<u8", nlist + 1)
    ix["arena_begin"] = arr("<u8", nlist + 1)
    ix["arena"] = arr("u1", int(ix["arena_begin"][-1]) + 64)
    ix["ref_arena_begin"] = arr("<u8", nlist + 1)
    ix["ref_arena"] = arr("u1", int(ix["ref_arena_begin"][-1]) + 64)
    ix["alpha"] = 1.0
    if off + 12 <= len(buf):
        lead, lam, tail = struct.unpack_from("<IfI", buf, off)
        if lead == OPT_MAGIC and tail == OPT_MAGIC and lam > 0 and lam == lam:
            ix["alpha"] = lam
    return ix
