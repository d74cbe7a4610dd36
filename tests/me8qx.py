"""Reader for the ME8QIDXX index file (reference me8q_io.cpp:29-176)."""
import numpy as np

MAGIC = b"ME8QIDXX"


def read(path):
    return parse(open(path, "rb").read())


def parse(buf):
    buf = bytes(buf)
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    if buf[:8] != MAGIC:
        raise ValueError("bad magic")
    off = 8
    n, dim, padded, nlist, bpv = (int(v) for v in take("<u8", 5))
    avg = float(take("<f4", 1)[0])
    flags = take("<i4", 4)
    nplan = int(take("<u8", 1)[0])
    plan = []
    for _ in range(nplan):
        ln = int(take("<u8", 1)[0])
        bits = int(take("<i4", 1)[0])
        plan.append((ln, bits))
    cent = take("<f4", nlist * padded).reshape(nlist, padded)
    ids = take("<u4", n)
    list_begin = take("<u8", nlist + 1)
    byte_begin = take("<u8", nlist + 1)
    codes = take("u1", int(byte_begin[-1]))
    rot_seed = int(take("<u8", 1)[0])
    nrot = int(take("<u8", 1)[0])
    signs = []
    for _ in range(nrot):
        nb = int(take("<u8", 1)[0])
        signs.append(take("u1", nb))
    assert off == len(buf), "trailing bytes"
    return dict(n=n, dim=dim, padded=padded, nlist=nlist, bytes_per_vector=bpv, avg_bits=avg,
                max_bits=int(flags[1]), drop_energy=int(flags[2]),
                prune_alpha=float(np.frombuffer(flags[3:4].tobytes(), "<f4")[0]), plan=plan, centroids=cent,
                ids=ids, list_begin=list_begin, byte_begin=byte_begin, codes=codes, rot_seed=rot_seed, signs=signs)
