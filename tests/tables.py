"""Static index tables (levels, refine table, scan class means) for feeding
the device library directly in tests, computed by the oracle."""
import numpy as np

import oracle_py as orc


def tables(bits, padded):
    l0, l1 = orc.branch_levels(bits, 0), orc.branch_levels(bits, 1)
    w = orc.class_means(bits)
    gm = np.empty(16, np.float64)
    for d in range(8):
        gm[2 * d] = 0.0 if (bits == 1 and d == 7) else w[0]
        gm[2 * d + 1] = 0.0 if (bits == 1 and d == 7) else w[1]
    stride = max(256, 1 << bits)
    rt = np.empty(2 * stride, np.float32)
    orc.lib().orc_refine_table(bits, padded, orc._p(rt))
    return dict(scale=orc.lib().orc_fixed_scale(bits), levels=np.concatenate([l0, l1]), gate_mid=gm,
                refine_tab=rt)


def layout(cluster_ids, nlist, bits, dim):
    """Cluster offsets, arena offsets and per-row slots (e8q_build.cpp:120-154, 225-231)."""
    P = (dim + 63) // 64 * 64
    blocks = P // 8
    counts = np.bincount(cluster_ids, minlength=nlist).astype(np.uint64)
    cb = np.zeros(nlist + 1, np.uint64)
    cb[1:] = np.cumsum(counts)
    bb = (128 + 32 * blocks + 63) // 64 * 64
    stride = blocks * (bits - 1) if bits >= 2 else 0
    rbb = (32 * stride + 256 + 63) // 64 * 64
    nb = (counts + 31) // 32
    ab = np.zeros(nlist + 1, np.uint64)
    ab[1:] = np.cumsum(nb * bb)
    rb = np.zeros(nlist + 1, np.uint64)
    rb[1:] = np.cumsum(nb * rbb)
    order = np.argsort(cluster_ids, kind="stable")
    slot = np.empty(len(cluster_ids), np.uint32)
    slot[order] = np.arange(len(cluster_ids), dtype=np.uint32)
    return dict(cluster_begin=cb, arena_begin=ab, ref_arena_begin=rb, slot=slot, padded=P)


def empty_arrays(n, dim, nlist, bits, lay, signs):
    P = lay["padded"]
    return dict(n=n, dim=dim, nlist=nlist, bits=bits, padded=P, alpha=1.0, signs=signs,
                cluster_begin=lay["cluster_begin"], arena_begin=lay["arena_begin"],
                ref_arena_begin=lay["ref_arena_begin"],
                ids=np.zeros(max(n, 1), np.uint32), mean_rot=np.zeros(P, np.float32),
                centroid_rot=np.zeros((nlist, P), np.float32), centroid_n2=np.zeros(nlist, np.float32),
                arena=np.zeros(int(lay["arena_begin"][-1]) + 64, np.uint8),
                ref_arena=np.zeros(int(lay["ref_arena_begin"][-1]) + 64, np.uint8))
