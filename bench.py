#!/usr/bin/env python3
"""QPS at recall@10 >= 0.95 on one MI355X, plain vs batched/served, with p50/p99 latency.

Default workload (the served ME8Q configuration): N = 10M unit-normalized
vectors, D = 1024, power-law spectrum (lambda_i ~ i^-1.1), 16384 lists with
Zipf(1.1) sizes (the largest ~1.47M vectors), taken to the PCA domain (fp64
covariance + eigh, descending) as the reference's runner does, ME8Q at
avg_bits 4 built on the device; 10k queries, k = 100, nprobe grid
8/16/32/64/128/256.  Synthetic data generated on the GPU from fixed seeds.

A step is one search call over the 10k-query batch (queries and outputs
resident in HBM).  For every search form and nprobe the script runs W warm-up
steps and K timed steps (synchronized on both sides; the plain form, whose
steps take seconds, one warm-up and --plain-steps timed steps), measures
recall@10 as the reference's recall_at_k (experiments/protocol.py:482-488:
overlap of the first 10 results with the exact fp64 top-10), and
interpolates QPS at the recall target on log(QPS) between the bracketing
grid points of the Pareto frontier (the reference's qps_at_recall rule,
experiments/plots/common.py:292-308).  Then a Poisson request stream (sizes
1..256, nprobe 8..256, k 10/100) goes through the continuous-batching server
(p50/p99 per request) and one plain search call per request.

Also reported: the batched form's dominant kernel by measured time with its
roofline (algorithmic bytes counted by the kernel / its launch time from HIP
events on the library's stream; HBM traffic from the committed rocprofv3
summary), and the reference CPU search (Me8qIndex::SearchIVFFastBatch on the
saved ME8QIDXX file) timed on this host on a bounded query sample.

Other workloads: `--workload e8q` (plain E8Q, 1M x 128 Gaussian mixture, 1024
lists, b = 4, k = 100, nprobe 16..128), `--workload me8q` (ME8Q 1M x 960,
4096 lists), `--workload zipf` (batched / MFMA E8Q, 10M x 768, 16384 Zipf
lists) and `--workload serve` (served E8Q or ME8Q at 1M).

One GPU only: --gpus N with N > 1 exits with a message.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "e8q-artifact_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("e8q", "me8q", "serve", "zipf", "me8q_zipf"), default="me8q_zipf",
                    help="zipf: the batched / MFMA E8Q configuration, 10M x 768, 16384 Zipf(1.1) lists, "
                         "k=10, per-query nprobe in {16,32,64,128}; me8q_zipf: the served ME8Q configuration, "
                         "10M x 1024 PCA-domain power law, 16384 Zipf(1.1) lists, avg_bits 4, plain vs batched "
                         "over an nprobe grid, then a Poisson request stream through the server")
    ap.add_argument("--serve-index", choices=("e8q", "me8q"), default="e8q")
    ap.add_argument("--serve-requests", type=int, default=1500)
    ap.add_argument("--serve-workers", type=int, default=3)
    ap.add_argument("--serve-rates", default="5,20,40,80",
                    help="offered request rates as multiples of the plain path's request rate")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=None, help="1M (e8q, me8q) / 10M (zipf)")
    ap.add_argument("--dim", type=int, default=None, help="128 (e8q) / 960 (me8q)")
    ap.add_argument("--nlist", type=int, default=None, help="1024 (e8q) / 4096 (me8q)")
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--k", type=int, default=None, help="100 (e8q, me8q) / 10 (zipf)")
    ap.add_argument("--bits", type=int, default=4)
    ap.add_argument("--grid", default=None, help="16,32,64,128 (e8q) / 32,64,128,256 (me8q)")
    ap.add_argument("--avg-bits", type=float, default=4.0)
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--modes", default=None,
                    help="plain,batched,batched_graph (e8q) / plain,batched,mfma,mfma_graph (zipf)")
    ap.add_argument("--latency-queries", type=int, default=200)
    ap.add_argument("--plain-steps", type=int, default=1, help="timed steps of the plain form (me8q_zipf)")
    ap.add_argument("--plain-budget", type=float, default=60.0,
                    help="me8q_zipf: seconds per plain step above which larger nprobe points are skipped")
    ap.add_argument("--serve-load", default="0.25,0.5,0.75",
                    help="me8q_zipf: offered request rates as fractions of the measured serving capacity")
    ap.add_argument("--serve-plain-requests", type=int, default=40,
                    help="me8q_zipf: requests of the stream timed through the plain path")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-sample", type=int, default=None, help="2000 queries (500 for me8q_zipf)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="only build and run the timed plain searches (for rocprofv3)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    a = ap.parse_args()
    me = a.workload == "me8q" or (a.workload == "serve" and a.serve_index == "me8q")
    zipf = a.workload == "zipf"
    mz = a.workload == "me8q_zipf"
    a.n = a.n or (10_000_000 if zipf or mz else 1_000_000)
    a.k = a.k or (10 if zipf else 100)
    a.modes = a.modes or ("plain,batched,mfma,mfma_graph" if zipf else "plain,batched" if a.workload == "me8q"
                          else "batched,plain" if mz else "plain,batched,batched_graph")
    a.dim = a.dim or (960 if me else 768 if zipf else 1024 if mz else 128)
    a.nlist = a.nlist or (4096 if me else 16384 if zipf or mz else 1024)
    a.grid = a.grid or ("32,64,128,256" if me else "8,16,32,64,128,256" if mz else "16,32,64,128")
    a.cpu_sample = a.cpu_sample or (500 if mz else 2000)
    return a


def copy_calibration(nbytes):
    """A plain copy of known size: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on it
    gives the counter-to-bytes factor used for the scan kernel's traffic."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 x {nbytes} bytes")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_data(torch, n, dim, nlist, nq, seed=42, center_sd=2.0):
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    centers = torch.randn(nlist, dim, generator=g, device=dev) * center_sd
    lab = torch.randint(0, nlist, (n,), generator=g, device=dev)
    base = centers[lab] + torch.randn(n, dim, generator=g, device=dev)
    qlab = torch.randint(0, nlist, (nq,), generator=g, device=dev)
    queries = centers[qlab] + torch.randn(nq, dim, generator=g, device=dev)
    from e8qlib import prep
    cid = prep.assign_ivf(base, centers).to(torch.int64)
    return base.contiguous(), centers.contiguous(), cid, queries.contiguous()


def ground_truth(torch, base, queries, k=10, cand=64):
    bn = (base.double() * base.double()).sum(1)
    out = torch.empty(queries.shape[0], k, dtype=torch.int64, device=base.device)
    for s in range(0, queries.shape[0], 512):
        q = queries[s:s + 512]
        d = bn.float()[None, :] - 2.0 * (q @ base.T)
        top = d.topk(cand, dim=1, largest=False).indices
        exact = ((base[top].double() - q[:, None, :].double()) ** 2).sum(2)
        order = exact.argsort(dim=1)[:, :k]
        out[s:s + 512] = torch.gather(top, 1, order)
    return out


def recall_at(gt, ids, k=10, within=None):
    """|exact top-k  intersect  first `within` results| / k, averaged (within=k: the
    reference's recall_at_k, experiments/protocol.py:482-488)."""
    w = k if within is None else within
    hits = 0
    for i in range(gt.shape[0]):
        hits += len(set(gt[i, :k].tolist()) & set(ids[i, :w].tolist()))
    return hits / float(k * gt.shape[0])


def pareto(rows):
    """The (recall, qps) points no other point beats on both."""
    pts = sorted({(r["recall"], r["qps"]) for r in rows if r["qps"] > 0 and r["recall"] >= 0},
                 key=lambda p: (-p[0], -p[1]))
    out, best_q = [], -1.0
    for rec, q in pts:   # by decreasing recall: keep a point only if it is faster than every point above it
        if q > best_q:
            out.append((rec, q))
            best_q = q
    return sorted(out)


def qps_at_recall(rows, target):
    """The reference's rule (experiments/plots/common.py:292-308: log-QPS
    interpolation between the points bracketing the target, NaN if the target
    is never reached) over the Pareto frontier of the grid, so that grid
    points tied on recall resolve to the fastest of them."""
    pts = pareto(rows)
    if not pts or max(r for r, _ in pts) < target:
        return float("nan")
    below = [p for p in pts if p[0] <= target]
    if below and below[-1][0] == target:
        return below[-1][1]
    lo = below[-1] if below else (pts[0][0] - 1e-9, pts[0][1])
    hi = min((p for p in pts if p[0] > target), key=lambda p: p[0])
    if lo[1] == hi[1]:
        return hi[1]
    t = (target - lo[0]) / max(hi[0] - lo[0], 1e-12)
    return math.exp(math.log(max(lo[1], 1e-9)) * (1 - t) + math.log(hi[1]) * t)


def cpu_baseline(idx, queries_np, k, nprobe, gpu_ids, args):
    """Reference e8qlib CPU search (oracle/_ref/ref_bench) on a bounded sample."""
    tool = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    sample = min(args.cpu_sample, queries_np.shape[0])
    import numpy as np
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "index.e8ax")
        idx.save(path)
        qf = os.path.join(tmp, "q.f32")
        queries_np[:sample].astype(np.float32).tofile(qf)
        out = os.path.join(tmp, "ids.u32")
        if os.path.exists(tool):
            kind = "reference"
            cmd = [tool, path, qf, str(sample), str(k), str(nprobe), str(args.cpu_threads), str(args.cpu_seconds),
                   out]
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stderr)
            res = json.loads(r.stdout.strip().splitlines()[-1])
            ids = np.fromfile(out, dtype=np.uint32).reshape(sample, k)
        else:
            kind = "port"
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import e8ax
            import oracle_py
            oi = oracle_py.OracleIndex.from_e8ax(e8ax.read(path))
            t0 = time.perf_counter()
            ids, _d, _s = oi.search(queries_np[:sample], k, nprobe, threads=args.cpu_threads)
            el = time.perf_counter() - t0
            res = {"qps_median": sample / el, "threads": args.cpu_threads, "seconds": el}
    same = float((ids == gpu_ids[:sample]).all(axis=1).mean())
    return {"value": round(res["qps_median"], 2), "unit": "queries/s", "cores": res["threads"], "kind": kind,
            "sample": f"{sample} of the {queries_np.shape[0]} queries, nprobe={nprobe}, k={k}, "
                      f"{res.get('rounds', 1)} rounds after 3 warm-ups, median round QPS "
                      f"({res['seconds']:.1f} s of CPU search)",
            "rows_identical_to_gpu_plain": same}


def make_me8q_data(torch, n, dim, nlist, nq, seed=42):
    """Power-law spectrum behind a seeded orthogonal rotation, then the PCA
    domain: base, queries and list centroids projected on the eigenvectors of
    the fp64 covariance in descending eigenvalue order; variance = the
    eigenvalues."""
    import numpy as np
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    lam = torch.arange(1, dim + 1, dtype=torch.float64, device=dev) ** -1.2
    rot, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((dim, dim)))
    rot = torch.from_numpy(rot).to(dev).float()
    sd = lam.sqrt().float()
    base = (torch.randn(n, dim, generator=g, device=dev) * sd) @ rot.T
    queries = (torch.randn(nq, dim, generator=g, device=dev) * sd) @ rot.T
    from e8qlib import prep
    pca = prep.pca_fit(base)
    prep.pca_transform(pca, base, out=base)
    queries = prep.pca_transform(pca, queries)
    var = pca.variance.astype(np.float32)
    pick = torch.randperm(n, generator=g, device=dev)[:nlist]
    centers = base[pick].clone()
    cid = prep.assign_ivf(base, centers).to(torch.int64)
    return base.contiguous(), centers.contiguous(), cid, queries.contiguous(), var


def me8q_layout(plan):
    """Scan record and refine row bytes of a plan (per 32-vector batch / vector)."""
    coded = [(ln, b) for ln, b in plan if b > 0]
    plane_cols = sum(ln // 8 for ln, _ in coded)
    rec_bytes = 32 * plane_cols + 160 * len(coded)
    refine_stride = 8 * len(coded) + sum((b - 1) * ln // 8 for ln, b in coded if b >= 2)
    return plane_cols, rec_bytes, refine_stride


def me8q_forms(idx, q_dev, nq, k, grid, modes, gt, args, plan, plain_budget_s=None):
    """Every form over the nprobe grid: W warm-up and K timed search_device
    calls per point (the plain form at the largest grid points, whose steps
    take seconds on head lists, with one warm-up at most)."""
    import numpy as np
    import torch
    plane_cols, rec_bytes, refine_stride = me8q_layout(plan)
    ids_dev = torch.empty((nq, k), dtype=torch.int32, device="cuda:0")
    d_dev = torch.empty((nq, k), dtype=torch.float32, device="cuda:0")
    results = {}
    for mode in modes:
        rows = []
        for nprobe in grid:
            steps, warm = args.steps, args.warmup
            if mode == "plain" and plain_budget_s is not None:
                steps, warm = max(1, min(steps, args.plain_steps)), min(warm, 1)
            log(f"me8q {mode} nprobe={nprobe}: {warm} warm-up + {steps} timed steps")
            for _ in range(warm):
                idx.search_device(q_dev.data_ptr(), nq, k, nprobe, ids_dev.data_ptr(), d_dev.data_ptr(), mode=mode)
            torch.cuda.synchronize()
            times, scan_ms, tot_ms, st = [], [], [], None
            kern = {f"refine_{ph}_{f}": 0.0 for ph in ("first", "cascade") for f in ("ms", "launches", "bytes")}
            for _ in range(steps):
                torch.cuda.synchronize()
                ts = time.perf_counter()
                st = idx.search_device(q_dev.data_ptr(), nq, k, nprobe, ids_dev.data_ptr(), d_dev.data_ptr(),
                                       mode=mode)
                torch.cuda.synchronize()
                times.append(time.perf_counter() - ts)
                scan_ms.append(st["device_scan_ms"])
                tot_ms.append(st["device_total_ms"])
                for f in kern:
                    kern[f] += st.get(f, 0) / steps
            ids = ids_dev.cpu().numpy().view(np.uint32)
            ms = 1000.0 * sum(times) / len(times)
            row = {"nprobe": nprobe, "ms_per_step": ms, "qps": nq / (ms / 1000.0),
                   "recall": recall_at(gt, ids, 10, 10) if gt is not None else -1.0,
                   "recall_10at100": recall_at(gt, ids, 10, k) if gt is not None else -1.0,
                   "scan_kernel_ms": sum(scan_ms) / len(scan_ms), "device_total_ms": sum(tot_ms) / len(tot_ms),
                   "batches": st["batches"], "survivors": st["survivors"], "candidates": st["candidates"],
                   "refined": st["refined"], "groups": st["groups"], "form": st["form"],
                   "overflowed": st["overflowed"], "pairs_scored": st["pairs_scored"], "items": st.get("items", 0)}
            if mode == "plain":
                algo = row["batches"] * rec_bytes + row["survivors"] * (refine_stride + 4)
                row["scan_algorithmic_bytes"] = algo
                row["scan_achieved_gbs"] = algo / (row["scan_kernel_ms"] / 1000.0) / 1e9
                extra = f"scan kernel {row['scan_kernel_ms']:.2f} ms = {row['scan_achieved_gbs']:.0f} GB/s"
            else:
                # int8 MFMA work: 32 one-hot entries per plane column per pair, 2 ops each
                ops = row["pairs_scored"] * 2 * 32 * plane_cols
                row["mfma_int8_ops"] = ops
                row["mfma_tops"] = ops / (row["scan_kernel_ms"] / 1000.0) / 1e12
                fill = row["pairs_scored"] / max(1, row["items"] * 32 * 256)
                row["item_fill"] = fill
                row.update(kern)
                extra = (f"MFMA scan {row['scan_kernel_ms']:.2f} ms = {row['mfma_tops']:.0f} TOP/s int8, "
                         f"{row['items']} items ({100 * fill:.1f}% of their 32 x 256 pairs scored), refine "
                         f"{kern['refine_first_ms']:.2f} ms (first runs, "
                         f"{kern['refine_first_bytes'] / max(kern['refine_first_ms'], 1e-9) / 1e6:.0f} GB/s) + "
                         f"{kern['refine_cascade_ms']:.2f} ms (cascades)")
            rows.append(row)
            log(f"me8q {mode} nprobe={nprobe}: {ms:.2f} ms/step, {row['qps']:.0f} QPS, recall@10 (10@10)="
                f"{row['recall']:.4f} (10@{k} {row['recall_10at100']:.4f}), {extra}, survivors "
                f"{row['survivors']}/{row['candidates']}, refined {row['refined']}")
            if mode == "plain" and plain_budget_s is not None and ms / 1000.0 > plain_budget_s:
                log(f"me8q plain: {ms / 1000.0:.1f} s per step, larger nprobe points skipped")
                break
        results[mode] = {"rows": rows, "qps_at_target": qps_at_recall(rows, args.target) if gt is not None else None}
    return results


def profile_traffic(kernel, workload_tag, nprobe):
    """HBM bytes per launch of `kernel` on this workload and nprobe, from the
    committed rocprofv3 summaries: a line
    `traffic: <kernel> <workload> nprobe=<n> bytes_per_launch=<B>` (FETCH_SIZE
    plus WRITE_SIZE per dispatch, scaled by the copy-kernel calibration the
    summary states).  None when no summary covers it."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            f = line.strip().strip("`").split()
            if len(f) == 5 and f[0] == "traffic:" and f[1] == kernel and f[2] == workload_tag \
                    and f[3] == f"nprobe={nprobe}" and f[4].startswith("bytes_per_launch="):
                return {"bytes_per_launch": float(f[4].split("=", 1)[1]),
                        "source": os.path.relpath(path, ROOT)}
    return None


def me8q_line(metric, workload, data, args, results, plan, bits_per_dim, k, nq, build_s, lat, cpu, extra_cfg,
              workload_tag=""):
    """The JSON line of an ME8Q workload: value = the best form's QPS at the
    recall target; roofline = the batched form's MFMA contraction (int8
    TOP/s) when it ran, else the plain scan's HBM rate."""
    plane_cols, rec_bytes, refine_stride = me8q_layout(plan)
    best_mode = max(results, key=lambda m: (results[m]["qps_at_target"] or 0.0))
    value = results[best_mode]["qps_at_target"]
    rows = results[best_mode]["rows"]
    op = next((r for r in rows if r["recall"] >= args.target), rows[-1])
    modes = {m: {"qps_at_target": (round(r["qps_at_target"], 1) if r["qps_at_target"] == r["qps_at_target"]
                                   and r["qps_at_target"] is not None else None),
                 "grid": [{kk: (round(v, 4) if isinstance(v, float) else v) for kk, v in row.items()}
                          for row in r["rows"]]} for m, r in results.items()}
    if "batched" in results:
        # the batched form's operating point: its smallest grid nprobe at the
        # recall target; the roofline is that of its kernel with the most time
        br = results["batched"]["rows"]
        rop = next((r for r in br if r["recall"] >= args.target), br[-1])
        kernels = {"M8MfmaScanKernel": rop["scan_kernel_ms"], "M8RefineKernel<true>": rop["refine_first_ms"],
                   "M8RefineKernel<false>": rop["refine_cascade_ms"]}
        dom = max(kernels, key=kernels.get)
        if dom == "M8MfmaScanKernel":
            roof = {"bound": "mfma", "kernel": "M8MfmaScanKernel (v_mfma_i32_32x32x32_i8)",
                    "achieved": round(rop["mfma_tops"], 1), "peak": 5000.0, "unit": "TOP/s",
                    "frac": round(rop["mfma_tops"] / 5000.0, 4), "traffic": None,
                    "basis": f"batched form at nprobe={rop['nprobe']}, one launch per search: {rop['pairs_scored']} "
                             f"(query, vector) pairs x 2 x 32 x {plane_cols} plane columns int8 ops / its time "
                             f"{rop['scan_kernel_ms']:.3f} ms (HIP events)"}
        else:
            ph = "first" if dom.endswith("<true>") else "cascade"
            nl = max(1, int(round(rop[f"refine_{ph}_launches"])))
            per_launch = rop[f"refine_{ph}_bytes"] / nl
            avg_ms = rop[f"refine_{ph}_ms"] / nl
            achieved = per_launch / (avg_ms / 1000.0) / 1e9
            traffic = profile_traffic(dom, workload_tag, rop["nprobe"])
            roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic["bytes_per_launch"] if traffic else None,
                    "basis": f"batched form at nprobe={rop['nprobe']}, {nl} launches per search: {per_launch:.4g} "
                             f"algorithmic bytes per launch (per candidate: 8-byte pool entry, {plane_cols}-byte "
                             f"plane row, 19 bytes of run header / banks per run, the fields of the runs it "
                             f"scores, its writes) / mean launch time {avg_ms:.3f} ms (HIP events around each "
                             f"launch on the library's stream)"}
            if traffic:
                roof["traffic_source"] = traffic["source"]
        roof["kernel_ms_per_search"] = {kk: round(vv, 3) for kk, vv in kernels.items()}
    else:
        pr = results["plain"]["rows"]
        rop = max(pr, key=lambda r: r["nprobe"])
        achieved = rop["scan_achieved_gbs"]
        traffic = profile_traffic("M8ScanKernel", workload_tag, rop["nprobe"])
        roof = {"bound": "hbm", "kernel": "M8ScanKernel", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "basis": f"per launch: {rop['batches']} scan records x {rec_bytes} B + {rop['survivors']} "
                         f"survivors x {refine_stride + 4} B, / mean kernel time {rop['scan_kernel_ms']:.3f} ms"}
        if traffic:
            roof["traffic_source"] = traffic["source"]
    line = {
        "metric": metric,
        "value": round(value, 1) if value is not None and value == value else None,
        "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000.0 * nq / value, 3) if value and value == value else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8 LUT / i32 coarse sums (int8 MFMA in the batched form), fp32 scores, fp64 bounds",
        "data": data,
        "config": dict({"workload": workload, "plan": plan, "bits_per_dim": bits_per_dim, "form": best_mode,
                        "operating_nprobe": op["nprobe"], "recall_at_operating_point": op["recall"],
                        "recall_definition": "recall@10 = |exact top-10 (fp64) & first 10 results| / 10, the "
                                             "reference's recall_at_k (experiments/protocol.py:482-488); "
                                             f"10@{k} per grid row as recall_10at100"}, **extra_cfg),
        "latency_us": lat,
        "modes": modes,
        "build_seconds": round(build_s, 2),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "plain" in results and "batched" in results:
        p, b = results["plain"]["qps_at_target"], results["batched"]["qps_at_target"]
        if p and b and p == p and b == b:
            line["batched_over_plain"] = round(b / p, 2)
    return line


def single_query_latency(idx, queries, k, nprobe, count, mode="plain"):
    lat = []
    qh = queries[:count].cpu().numpy()
    for i in range(qh.shape[0]):
        ts = time.perf_counter()
        idx.search(qh[i:i + 1], k, nprobe, mode=mode)
        lat.append((time.perf_counter() - ts) * 1e6)
    lat.sort()
    return {"p50": round(lat[len(lat) // 2], 1), "p99": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 1),
            "what": f"single-query search() calls ({mode} form, nprobe={nprobe})"}


def run_me8q(args):
    import numpy as np
    import torch
    import e8qlib

    grid = [int(x) for x in args.grid.split(",")]
    modes = [m for m in args.modes.split(",") if m]
    t0 = time.perf_counter()
    base, centers, cid, queries, var = make_me8q_data(torch, args.n, args.dim, args.nlist, args.nq)
    gt = None if args.profile else ground_truth(torch, base, queries).cpu().numpy()
    torch.cuda.synchronize()
    log(f"data + ground truth: {time.perf_counter() - t0:.1f} s")
    idx = e8qlib.Me8qIndex(args.dim, args.avg_bits, 0)
    tb = time.perf_counter()
    idx.build(base.cpu().numpy(), centers.cpu().numpy(), cid.to(torch.int32).cpu().numpy().astype(np.uint32),
              variance=var)
    build_s = time.perf_counter() - tb
    plan = [tuple(p) for p in idx.plan]
    log(f"build: {build_s:.2f} s, plan {plan}, {idx.bits_per_dim:.3f} bits/dim")
    del base
    torch.cuda.empty_cache()
    nq, k = args.nq, args.k
    q_dev = queries.contiguous()
    results = me8q_forms(idx, q_dev, nq, k, grid, modes, gt, args, plan)
    if args.profile:
        return
    best = max(results, key=lambda m: (results[m]["qps_at_target"] or 0.0))
    rows = results[best]["rows"]
    op = next((r for r in rows if r["recall"] >= args.target), rows[-1])
    lat = single_query_latency(idx, queries, k, op["nprobe"], args.latency_queries, best)
    cpu = None
    if not args.no_cpu and "plain" in results:
        prow = results["plain"]["rows"]
        opp = next((r for r in prow if r["recall"] >= args.target), prow[-1])
        ids_dev = torch.empty((nq, k), dtype=torch.int32, device="cuda:0")
        d_dev = torch.empty((nq, k), dtype=torch.float32, device="cuda:0")
        idx.search_device(q_dev.data_ptr(), nq, k, opp["nprobe"], ids_dev.data_ptr(), d_dev.data_ptr())
        torch.cuda.synchronize()
        cpu = cpu_baseline_me8q(idx, queries.cpu().numpy(), k, opp["nprobe"],
                                ids_dev.cpu().numpy().view(np.uint32), args)
        log(f"cpu baseline: {cpu}")
    line = me8q_line("QPS at recall@10>=0.95 (ME8Q mixed-rate search, plain vs batched, 1 GPU); p50/p99 latency",
                     f"ME8Q mixed-rate scan: N={args.n}, D={args.dim} PCA-domain power law, {args.nlist} lists, "
                     f"avg_bits={args.avg_bits}, {nq} queries, k={k}, nprobe {args.grid}",
                     "synthetic power-law spectrum (lambda_i ~ i^-1.2), seeded rotation, PCA domain, generated on the GPU",
                     args, results, plan, idx.bits_per_dim, k, nq, build_s, lat, cpu, {},
                     workload_tag=f"me8q_{args.n // 1_000_000}Mx{args.dim}")
    print(json.dumps(line), flush=True)


def cpu_baseline_me8q(idx, queries_np, k, nprobe, gpu_ids, args):
    """Reference e8qlib ME8Q CPU search (oracle/_ref/ref_bench, SearchIVFFastBatch) on a bounded sample."""
    import numpy as np
    tool = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if not os.path.exists(tool):
        return None
    sample = min(args.cpu_sample, queries_np.shape[0])
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "index.me8q")
        idx.save(path)
        qf = os.path.join(tmp, "q.f32")
        queries_np[:sample].astype(np.float32).tofile(qf)
        out = os.path.join(tmp, "ids.u32")
        cmd = [tool, path, qf, str(sample), str(k), str(nprobe), str(args.cpu_threads), str(args.cpu_seconds), out]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        res = json.loads(r.stdout.strip().splitlines()[-1])
        ids = np.fromfile(out, dtype=np.uint32).reshape(sample, k)
    same = float((ids == gpu_ids[:sample]).all(axis=1).mean())
    return {"value": round(res["qps_median"], 2), "unit": "queries/s", "cores": res["threads"], "kind": "reference",
            "sample": f"{sample} of the {queries_np.shape[0]} queries, nprobe={nprobe}, k={k}, "
                      f"{res.get('rounds', 1)} rounds after 3 warm-ups, median round QPS "
                      f"({res['seconds']:.1f} s of CPU search; Me8qIndex::SearchIVFFastBatch on the saved "
                      f"ME8QIDXX file, loaded in {res.get('load_seconds', 0.0):.1f} s)",
            "rows_identical_to_gpu_plain": same}


def run_serve(args):
    """Continuous-batching serving (e8qlib.Server) against the plain path (one
    search() call per request, back to back) on the same request stream:
    sizes log-uniform in [1, 256] queries, nprobe log-uniform over
    {8, ..., 256}, k uniform over {10, 100}, Poisson arrivals."""
    import threading

    import numpy as np
    import torch
    import e8qlib

    rng = np.random.default_rng(7)
    t0 = time.perf_counter()
    if args.serve_index == "me8q":
        base, centers, cid, queries, var = make_me8q_data(torch, args.n, args.dim, args.nlist, args.nq)
        idx = e8qlib.Me8qIndex(args.dim, args.avg_bits, 0)
        idx.build(base.cpu().numpy(), centers.cpu().numpy(), cid.to(torch.int32).cpu().numpy().astype(np.uint32),
                  variance=var)
        workload = f"ME8Q N={args.n} D={args.dim} {args.nlist} lists avg_bits={args.avg_bits}"
    else:
        base, centers, cid, queries = make_data(torch, args.n, args.dim, args.nlist, args.nq)
        idx = e8qlib.E8qIndex(args.dim, args.bits, 42)
        idx.build(base.cpu().numpy(), centers.cpu().numpy(), cid.to(torch.int32).cpu().numpy().astype(np.uint32))
        workload = f"E8Q N={args.n} D={args.dim} {args.nlist} lists (centers N(0,4I)) b={args.bits}"
    log(f"index: {workload}, {time.perf_counter() - t0:.1f} s")
    qpool = queries.cpu().numpy()
    del base
    nprobes = [p for p in (8, 16, 32, 64, 128, 256) if p <= args.nlist]
    reqs = []
    for _ in range(args.serve_requests):
        nq = int(np.exp(rng.uniform(0.0, np.log(256.0))))
        rows = rng.integers(0, qpool.shape[0], nq)
        reqs.append((np.ascontiguousarray(qpool[rows]), int(rng.choice([10, 100])), int(rng.choice(nprobes))))
    total_q = sum(r[0].shape[0] for r in reqs)
    # plain path: one blocking search per request
    for q, k, npb in reqs[:20]:
        idx.search(q, k, npb)
    ts = time.perf_counter()
    plain_lat = []
    for q, k, npb in reqs:
        a = time.perf_counter()
        idx.search(q, k, npb)
        plain_lat.append((time.perf_counter() - a) * 1e6)
    plain_s = time.perf_counter() - ts
    plain_rps = len(reqs) / plain_s
    plain_qps = total_q / plain_s
    plain_lat.sort()
    log(f"plain: {plain_qps:.0f} QPS ({plain_rps:.0f} requests/s), service p50 {plain_lat[len(plain_lat) // 2]:.0f} us")
    srv = e8qlib.Server(idx, workers=args.serve_workers, max_batch=8192)
    warm = [srv.submit(q, k, npb) for q, k, npb in reqs[:50]]
    for t in warm:
        srv.wait(t)
    runs = []
    for mult in [float(x) for x in args.serve_rates.split(",")]:
        rate = plain_rps * mult
        gaps = rng.exponential(1.0 / rate, len(reqs))
        tickets = [None] * len(reqs)
        lat = [0.0] * len(reqs)
        done_t = [0.0] * len(reqs)

        def waiter():
            for i in range(len(reqs)):
                while tickets[i] is None:
                    time.sleep(0.0002)
                _, _, info = srv.wait(tickets[i])
                lat[i] = info["latency_us"]
                done_t[i] = time.perf_counter()

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        nxt = start
        for i, (q, k, npb) in enumerate(reqs):
            nxt += gaps[i]
            d = nxt - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(q, k, npb)
        th.join()
        span = max(done_t) - start
        ls = sorted(lat)
        run = {"offered_rps": round(rate, 1), "offered_x_plain": mult, "qps": round(total_q / span, 1),
               "requests_per_s": round(len(reqs) / span, 1), "p50_us": round(ls[len(ls) // 2], 1),
               "p99_us": round(ls[min(len(ls) - 1, int(0.99 * len(ls)))], 1)}
        runs.append(run)
        log(f"served at {mult}x plain rate: {run['qps']:.0f} QPS, p50 {run['p50_us']:.0f} us, p99 {run['p99_us']:.0f} us")
    st = srv.stats()
    srv.close()
    best = max(runs, key=lambda r: r["qps"])
    line = {
        "metric": "served QPS with p50/p99 latency (continuous batching vs one search call per request, 1 GPU)",
        "value": best["qps"], "unit": "queries/s", "n_gpus": 1, "steps": len(runs), "warmup": 1,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "same as the index's plain search", "data": "synthetic, generated on the GPU",
        "config": {"workload": f"served {workload}: {len(reqs)} requests ({total_q} queries), sizes log-uniform "
                               f"[1,256], nprobe log-uniform {nprobes}, k in {{10,100}}, Poisson arrivals",
                   "workers": args.serve_workers},
        "plain": {"qps": round(plain_qps, 1), "requests_per_s": round(plain_rps, 1),
                  "service_p50_us": round(plain_lat[len(plain_lat) // 2], 1),
                  "service_p99_us": round(plain_lat[min(len(plain_lat) - 1, int(0.99 * len(plain_lat)))], 1)},
        "served_runs": runs, "served_over_plain": round(best["qps"] / plain_qps, 2), "server_stats": st,
        "latency_us": {"p50": best["p50_us"], "p99": best["p99_us"], "what": "per request, submit to results"},
    }
    print(json.dumps(line), flush=True)


def make_zipf_data(torch, n, dim, nlist, nq, seed=7, s=1.1, spectrum=1.0):
    """Unit-normalized anisotropic vectors in Zipf-sized lists: list sizes
    proportional to c^-s over nlist lists (summing to n); per-dimension
    standard deviations sqrt(lambda_i), lambda_i = i^-spectrum; centers ~
    N(0, 4 Lambda), points = center + N(0, Lambda), then scaled to unit norm;
    cluster_ids = the generating list, centroids = per-list means; queries
    drawn like the points (list chosen in proportion to its size)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    w = torch.arange(1, nlist + 1, dtype=torch.float64) ** (-s)
    sizes = torch.floor(w / w.sum() * n).to(torch.int64)
    sizes[:n - int(sizes.sum())] += 1
    sd = (torch.arange(1, dim + 1, dtype=torch.float64, device=dev) ** (-spectrum)).sqrt().float()
    centers = torch.randn(nlist, dim, generator=g, device=dev) * (2.0 * sd)
    lab = torch.repeat_interleave(torch.arange(nlist, device=dev), sizes.to(dev))
    base = torch.empty(n, dim, device=dev)
    for st in range(0, n, 1 << 20):
        e = min(n, st + (1 << 20))
        x = centers[lab[st:e]] + torch.randn(e - st, dim, generator=g, device=dev) * sd
        base[st:e] = x / x.norm(dim=1, keepdim=True)
    cent = torch.zeros(nlist, dim, device=dev, dtype=torch.float64)
    # lists are contiguous row ranges, largest first: the large ones summed
    # directly (an index_add_ into one row serializes on its atomics), the
    # rest with index_add_
    off = torch.cat([torch.zeros(1, dtype=torch.int64), sizes.cumsum(0)]).tolist()
    big = int((sizes >= 65536).sum())
    for c in range(big):
        for st in range(off[c], off[c + 1], 1 << 20):
            cent[c] += base[st:min(off[c + 1], st + (1 << 20))].double().sum(0)
    for st in range(off[big], n, 1 << 20):
        e = min(n, st + (1 << 20))
        cent.index_add_(0, lab[st:e], base[st:e].double())
    cent = (cent / sizes.to(dev).clamp(min=1)[:, None].double()).float()
    pick = torch.randint(0, n, (nq,), generator=g, device=dev)
    queries = centers[lab[pick]] + torch.randn(nq, dim, generator=g, device=dev) * sd
    queries = queries / queries.norm(dim=1, keepdim=True)
    return base, cent.contiguous(), lab.to(torch.int32), queries.contiguous(), sizes


def ground_truth_chunked(torch, base, queries, k=10, cand=64, qchunk=1024, bchunk=1 << 21):
    """Exact top-k: fp32 GEMM candidates per base chunk, merged, re-ranked in fp64."""
    n = base.shape[0]
    bn = torch.empty(n, device=base.device)
    for st in range(0, n, bchunk):
        x = base[st:st + bchunk].double()
        bn[st:st + bchunk] = (x * x).sum(1).float()
    out = torch.empty(queries.shape[0], k, dtype=torch.int64, device=base.device)
    for qs in range(0, queries.shape[0], qchunk):
        q = queries[qs:qs + qchunk]
        best_d, best_i = None, None
        for st in range(0, n, bchunk):
            d = bn[None, st:st + bchunk] - 2.0 * (q @ base[st:st + bchunk].T)
            td, ti = d.topk(cand, dim=1, largest=False)
            ti = ti + st
            if best_d is None:
                best_d, best_i = td, ti
            else:
                md = torch.cat([best_d, td], 1)
                mi = torch.cat([best_i, ti], 1)
                sel = md.topk(cand, dim=1, largest=False).indices
                best_d, best_i = torch.gather(md, 1, sel), torch.gather(mi, 1, sel)
        exact = ((base[best_i].double() - q[:, None, :].double()) ** 2).sum(2)
        order = exact.argsort(dim=1)[:, :k]
        out[qs:qs + qchunk] = torch.gather(best_i, 1, order)
    return out


def run_zipf(args):
    """The batched / MFMA configuration: every form over one 10k-query batch with
    per-query nprobe uniform in {16, 32, 64, 128}."""
    import numpy as np
    import torch
    import e8qlib

    t0 = time.perf_counter()
    base, cent, lab, queries, sizes = make_zipf_data(torch, args.n, args.dim, args.nlist, args.nq)
    torch.cuda.synchronize()
    log(f"zipf data: {time.perf_counter() - t0:.1f} s, largest list {int(sizes.max())}, smallest {int(sizes.min())}")
    t0 = time.perf_counter()
    gt = ground_truth_chunked(torch, base, queries, k=10).cpu().numpy()
    torch.cuda.synchronize()
    log(f"ground truth: {time.perf_counter() - t0:.1f} s")
    idx = e8qlib.E8qIndex(args.dim, args.bits, 42)
    lab_np = lab.cpu().numpy().astype(np.uint32)
    tb = time.perf_counter()
    idx.build_device(base.data_ptr(), cent.data_ptr(), args.nlist, lab_np, lab.data_ptr())
    torch.cuda.synchronize()
    build_s = time.perf_counter() - tb
    log(f"build: {build_s:.2f} s ({args.n / build_s / 1e6:.2f} M vectors/s, base in HBM)")
    del base
    torch.cuda.empty_cache()
    nq, k = args.nq, args.k
    rng = np.random.default_rng(11)
    npq = rng.choice(np.array([16, 32, 64, 128], np.uint32), size=nq)
    q_dev = queries.contiguous()
    ids_dev = torch.empty((nq, k), dtype=torch.int32, device="cuda:0")
    d_dev = torch.empty((nq, k), dtype=torch.float32, device="cuda:0")
    P = (args.dim + 63) // 64 * 64
    K = P * 16
    results = {}
    for mode in [m for m in args.modes.split(",") if m]:
        form = mode[:-len("_graph")] if mode.endswith("_graph") else mode
        e8qlib.set_search_graphs(mode != form)
        for _ in range(max(args.warmup, 2 if mode != form else 0)):
            idx.search_device(q_dev.data_ptr(), nq, k, 128, ids_dev.data_ptr(), d_dev.data_ptr(), form, npq)
        torch.cuda.synchronize()
        times, scan_ms, st = [], [], None
        for _ in range(args.steps):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            st = idx.search_device(q_dev.data_ptr(), nq, k, 128, ids_dev.data_ptr(), d_dev.data_ptr(), form, npq)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - ts)
            scan_ms.append(st["device_scan_ms"])
        if mode != form:
            info = e8qlib.search_context_info()
            e8qlib.set_search_graphs(False)
        ids = ids_dev.cpu().numpy().view(np.uint32)
        ms = 1000.0 * sum(times) / len(times)
        row = {"ms_per_step": ms, "qps": nq / (ms / 1000.0), "recall": recall_at(gt, ids, 10, 10),
               "scan_kernel_ms": sum(scan_ms) / len(scan_ms), "candidates": st["candidates"],
               "survivors": st["survivors"], "batches": st["batches"], "records_read": st["scanned"]}
        if form == "mfma" and mode == form:
            flops = st["candidates"] * 4.0 * K
            row["contraction_tflops"] = flops / (row["scan_kernel_ms"] / 1000.0) / 1e12
            row["useful_tflops"] = st["candidates"] * 2.0 * P / (row["scan_kernel_ms"] / 1000.0) / 1e12
        if mode != form:
            row["graph"] = info
        results[mode] = row
        log(f"{mode}: {ms:.2f} ms/step, {row['qps']:.0f} QPS, recall@10 {row['recall']:.4f}, "
            f"scan kernel {row['scan_kernel_ms']:.2f} ms, candidates {row['candidates']}, "
            f"survivors {row['survivors']}"
            + (f", contraction {row['contraction_tflops']:.0f} TFLOP/s" if "contraction_tflops" in row else ""))
    ok = {m: r for m, r in results.items() if r["recall"] >= args.target}
    best = max(ok, key=lambda m: ok[m]["qps"]) if ok else max(results, key=lambda m: results[m]["recall"])
    plain_qps = results.get("plain", {}).get("qps")
    mf = results.get("mfma")
    line = {
        "metric": "QPS at recall@10>=0.95 (plain vs batched/MFMA, 1 GPU)",
        "value": round(results[best]["qps"], 1) if ok else None,
        "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(results[best]["ms_per_step"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 hi+lo MFMA / fp32 accumulate (mfma); u8 / i32 / fp32 (plain, batched)",
        "data": "synthetic unit-normalized anisotropic mixture (lambda_i ~ i^-1), Zipf(1.1) list sizes, generated on the GPU (seed 7)",
        "config": {"workload": f"batched MFMA E8Q: N={args.n}, D={args.dim}, {args.nlist} Zipf lists, b={args.bits}, "
                               f"{nq} queries, k={k}, nprobe per query uniform in {{16,32,64,128}}",
                   "form": best, "largest_list": int(sizes.max()), "smallest_list": int(sizes.min())},
        "speedup_vs_plain": round(results[best]["qps"] / plain_qps, 2) if plain_qps and ok else None,
        "forms": {m: {kk: (round(v, 4) if isinstance(v, float) else v) for kk, v in r.items()}
                  for m, r in results.items()},
        "build_seconds": round(build_s, 2),
        "roofline": ({"bound": "mfma", "kernel": "MfmaScanKernel", "achieved": round(mf["contraction_tflops"], 1),
                      "peak": 2500.0, "unit": "TFLOP/s", "frac": round(mf["contraction_tflops"] / 2500.0, 4),
                      "traffic": None,
                      "basis": f"per launch {mf['candidates']} (query, vector) pairs x 4 K flops, K = {P} x 16 "
                               f"(one-hot) and x2 for hi+lo; useful 2 P flops per pair: "
                               f"{mf['useful_tflops']:.2f} TFLOP/s; / mean kernel time {mf['scan_kernel_ms']:.3f} ms"}
                     if mf else None),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def pca_domain(torch, base, queries, cent):
    """The reference's PCA step (e8qlib.prep: fp64 covariance, eigh,
    descending, fp64 transform) applied in place to base and to queries /
    centroids; returns the eigenvalues (the per-dimension variance)."""
    import numpy as np
    from e8qlib import prep
    pca = prep.pca_fit(base)
    prep.pca_transform(pca, base, out=base)
    return (prep.pca_transform(pca, queries).contiguous(), prep.pca_transform(pca, cent).contiguous(),
            pca.variance.astype(np.float32))


def serve_stream(rng, qpool, count, nprobes):
    """Requests: sizes log-uniform in [1, 256], nprobe uniform over the
    log-spaced grid, k uniform over {10, 100}."""
    import numpy as np
    reqs = []
    for _ in range(count):
        nq = int(np.exp(rng.uniform(0.0, np.log(256.0))))
        rows = rng.integers(0, qpool.shape[0], nq)
        reqs.append((np.ascontiguousarray(qpool[rows]), int(rng.choice([10, 100])), int(rng.choice(nprobes))))
    return reqs


def serve_run(srv, reqs, rate, rng):
    """Submits the stream with exponential gaps at `rate` requests/s (rate
    None: all at once) and waits; QPS over the span, p50/p99 per request.
    Logs a progress line every 200 submitted requests."""
    import threading
    tickets = [None] * len(reqs)
    lat = [0.0] * len(reqs)
    done_t = [0.0] * len(reqs)
    gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else [0.0] * len(reqs)

    def waiter():
        for i in range(len(reqs)):
            while tickets[i] is None:
                time.sleep(0.0002)
            _, _, info = srv.wait(tickets[i])
            lat[i] = info["latency_us"]
            done_t[i] = time.perf_counter()

    th = threading.Thread(target=waiter)
    th.start()
    start = time.perf_counter()
    nxt = start
    for i, (q, k, npb) in enumerate(reqs):
        nxt += gaps[i]
        d = nxt - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets[i] = srv.submit(q, k, npb)
        if (i + 1) % 200 == 0:
            log(f"  submitted {i + 1}/{len(reqs)} requests")
    th.join()
    span = max(done_t) - start
    ls = sorted(lat)
    total_q = sum(r[0].shape[0] for r in reqs)
    return {"offered_rps": round(rate, 1) if rate else "saturation", "qps": round(total_q / span, 1),
            "requests_per_s": round(len(reqs) / span, 1), "p50_us": round(ls[len(ls) // 2], 1),
            "p99_us": round(ls[min(len(ls) - 1, int(0.99 * len(ls)))], 1)}


def run_me8q_zipf(args):
    """The served ME8Q configuration: 10M x 1024 unit-normalized vectors with a
    power-law spectrum (lambda_i ~ i^-1.1) in 16384 Zipf(1.1)-sized lists,
    taken to the PCA domain, ME8Q at avg_bits 4 built on the device; plain vs
    batched QPS at recall over an nprobe grid on a 10k-query batch, then a
    Poisson request stream (sizes log-uniform in [1, 256], nprobe over
    {8..256}, k in {10, 100}) through the continuous-batching server against
    one plain search call per request."""
    import numpy as np
    import torch
    import e8qlib

    grid = [int(x) for x in args.grid.split(",")]
    modes = [m for m in args.modes.split(",") if m]
    t0 = time.perf_counter()
    base, cent, lab, queries, sizes = make_zipf_data(torch, args.n, args.dim, args.nlist, args.nq, spectrum=1.1)
    torch.cuda.synchronize()
    log(f"zipf data: {time.perf_counter() - t0:.1f} s, largest list {int(sizes.max())}, smallest {int(sizes.min())}")
    t0 = time.perf_counter()
    queries, cent, var = pca_domain(torch, base, queries, cent)
    torch.cuda.synchronize()
    log(f"PCA domain: {time.perf_counter() - t0:.1f} s, variance head {var[:4]}")
    gt = None
    if not args.profile:
        t0 = time.perf_counter()
        gt = ground_truth_chunked(torch, base, queries, k=10).cpu().numpy()
        torch.cuda.synchronize()
        log(f"ground truth: {time.perf_counter() - t0:.1f} s")
    idx = e8qlib.Me8qIndex(args.dim, args.avg_bits, 0)
    lab_dev = lab.contiguous()
    tb = time.perf_counter()
    idx.build_device(base.data_ptr(), cent.cpu().numpy(), lab_dev.cpu().numpy().view(np.uint32), lab_dev.data_ptr(),
                     variance=var)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - tb
    plan = [tuple(p) for p in idx.plan]
    log(f"build: {build_s:.2f} s, plan {plan}, {idx.bits_per_dim:.3f} bits/dim, {idx.device_bytes / 1e9:.2f} GB on device")
    del base, lab_dev
    torch.cuda.empty_cache()
    nq, k = args.nq, args.k
    q_dev = queries.contiguous()
    results = me8q_forms(idx, q_dev, nq, k, grid, modes, gt, args, plan, plain_budget_s=args.plain_budget)
    if args.profile:
        return
    # serving
    rng = np.random.default_rng(7)
    qpool = queries.cpu().numpy()
    nprobes = [p for p in (8, 16, 32, 64, 128, 256) if p <= args.nlist]
    reqs = serve_stream(rng, qpool, args.serve_requests, nprobes)
    total_q = sum(r[0].shape[0] for r in reqs)
    sample = reqs[:args.serve_plain_requests]
    log(f"serving: {len(reqs)} requests, {total_q} queries; plain path on the first {len(sample)}")
    idx.search(sample[0][0], sample[0][1], sample[0][2])
    plain_lat = []
    ts = time.perf_counter()
    timed = 0
    for q, kk, npb in sample:
        a = time.perf_counter()
        idx.search(q, kk, npb)
        plain_lat.append((time.perf_counter() - a) * 1e6)
        timed += 1
        log(f"plain request {timed}: {q.shape[0]} queries, k={kk}, nprobe={npb}, {plain_lat[-1] / 1e3:.1f} ms")
        if time.perf_counter() - ts > args.plain_budget and timed >= 3:
            break
    sample = sample[:timed]
    plain_s = time.perf_counter() - ts
    plain_rps = len(sample) / plain_s
    plain_qps = sum(r[0].shape[0] for r in sample) / plain_s
    plain_lat.sort()
    log(f"plain: {plain_qps:.0f} QPS ({plain_rps:.2f} requests/s), service p50 {plain_lat[len(plain_lat) // 2]:.0f} us")
    # each worker reserves an equal share of half the free HBM: the batched
    # form's candidate pool takes what its batch leaves of the reservation,
    # and a head-list query at nprobe 256 needs far more than the per-query
    # default the minimal reservation assumes (queries past the pool would
    # be answered by the plain form; server_stats.overflowed counts them)
    free_b, _total_b = torch.cuda.mem_get_info()
    ws = int(0.5 * free_b / args.serve_workers)
    srv = e8qlib.Server(idx, workers=args.serve_workers, max_batch=4096, mode="batched", workspace_bytes=ws)
    log(f"server: {args.serve_workers} workers, {ws / 1e9:.1f} GB workspace each")
    # warm-up (untimed): the batched form tunes its grids once per batch
    # shape (query count, nprobe and k buckets) and the workers share the
    # table, so the stream is replayed in groups that form batches of about
    # 64 .. 8192 queries, then all at once
    pos = 0
    for target in (64, 128, 256, 512, 1024, 2048, 4096, 8192):
        group, got = [], 0
        while got < target:
            q, kk, npb = reqs[pos % len(reqs)]
            group.append(srv.submit(q, kk, npb))
            got += q.shape[0]
            pos += 1
        for t in group:
            srv.wait(t)
    for t in [srv.submit(q, kk, npb) for q, kk, npb in reqs]:
        srv.wait(t)
    log(f"server warmed up ({srv.stats()['batches']} batches)")
    # capacity first (every request submitted at once), then Poisson streams
    # offered at fractions of it, each bounded to ~20 s of offered traffic
    sat = serve_run(srv, reqs, None, rng)
    sat["offered"] = "saturation"
    log(f"served (saturation): {sat['qps']:.0f} QPS, {sat['requests_per_s']:.1f} requests/s, "
        f"p50 {sat['p50_us']:.0f} us, p99 {sat['p99_us']:.0f} us")
    runs = [sat]
    for frac in [float(x) for x in args.serve_load.split(",")]:
        rate = sat["requests_per_s"] * frac
        sub = reqs[:max(50, min(len(reqs), int(rate * 20.0)))]
        run = serve_run(srv, sub, rate, rng)
        run["offered"] = f"{frac:g} x capacity"
        run["offered_x_plain"] = round(rate / plain_rps, 1)
        runs.append(run)
        log(f"served ({run['offered']}, {run['offered_x_plain']}x plain): {run['qps']:.0f} QPS, "
            f"{run['requests_per_s']:.1f} requests/s, p50 {run['p50_us']:.0f} us, p99 {run['p99_us']:.0f} us")
    sst = srv.stats()
    srv.close()
    best = max(runs, key=lambda r: r["qps"])
    lat_run = runs[-1] if len(runs) > 1 else runs[0]   # the highest offered fraction below capacity
    serving = {"requests": len(reqs), "queries": total_q, "workers": args.serve_workers,
               "stream": "sizes log-uniform [1,256], nprobe uniform over " + str(nprobes) +
                         ", k in {10,100}, Poisson arrivals (and one run with every request submitted at once)",
               "plain": {"qps": round(plain_qps, 1), "requests_per_s": round(plain_rps, 3),
                         "requests_timed": len(sample),
                         "service_p50_us": round(plain_lat[len(plain_lat) // 2], 1),
                         "service_p99_us": round(plain_lat[min(len(plain_lat) - 1, int(0.99 * len(plain_lat)))], 1)},
               "served_runs": runs, "served_over_plain": round(best["qps"] / plain_qps, 2), "server_stats": sst}
    # the reference CPU search on the same index (saved as an ME8QIDXX file),
    # at the plain form's operating nprobe, on a bounded query sample
    cpu = None
    if not args.no_cpu and "plain" in results:
        prow = results["plain"]["rows"]
        opp = next((r for r in prow if r["recall"] >= args.target), prow[-1])
        ids_dev = torch.empty((nq, k), dtype=torch.int32, device="cuda:0")
        d_dev = torch.empty((nq, k), dtype=torch.float32, device="cuda:0")
        idx.search_device(q_dev.data_ptr(), nq, k, opp["nprobe"], ids_dev.data_ptr(), d_dev.data_ptr(), mode="plain")
        torch.cuda.synchronize()
        log(f"cpu baseline: reference ME8Q search of {args.cpu_sample} queries at nprobe={opp['nprobe']}, "
            f"{args.cpu_threads} threads")
        cpu = cpu_baseline_me8q(idx, qpool, k, opp["nprobe"], ids_dev.cpu().numpy().view(np.uint32), args)
        log(f"cpu baseline: {cpu}")
    line = me8q_line("QPS at recall@10>=0.95 (ME8Q, plain vs batched/served, 1 GPU); p50/p99 latency",
                     f"served ME8Q: N={args.n}, D={args.dim} unit-normalized power law (lambda_i ~ i^-1.1) in the "
                     f"PCA domain, {args.nlist} Zipf(1.1) lists, avg_bits={args.avg_bits}, {nq} queries, k={k}, "
                     f"nprobe {args.grid}",
                     "synthetic, generated on the GPU (seed 7); PCA by fp64 covariance + eigh",
                     args, results, plan, idx.bits_per_dim, k, nq, build_s,
                     {"p50": lat_run["p50_us"], "p99": lat_run["p99_us"],
                      "what": f"per request through the server, Poisson arrivals at {lat_run['offered']} "
                              f"({lat_run['qps']:.0f} QPS served)"},
                     cpu, {"list_sizes": {"largest": int(sizes.max()), "smallest": int(sizes.min())}},
                     workload_tag=f"me8q_zipf_{args.n // 1_000_000}Mx{args.dim}")
    line["serving"] = serving
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "me8q":
        run_me8q(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return
    if args.workload == "zipf":
        run_zipf(args)
        return
    if args.workload == "me8q_zipf":
        run_me8q_zipf(args)
        return
    import numpy as np
    import torch
    import e8qlib

    grid = [int(x) for x in args.grid.split(",")]
    modes = [m for m in args.modes.split(",") if m]
    t0 = time.perf_counter()
    base, centers, cid, queries = make_data(torch, args.n, args.dim, args.nlist, args.nq)
    gt = None if args.profile else ground_truth(torch, base, queries).cpu().numpy()
    torch.cuda.synchronize()
    log(f"data + ground truth: {time.perf_counter() - t0:.1f} s")
    idx = e8qlib.E8qIndex(args.dim, args.bits, 42)
    tb = time.perf_counter()
    idx.build(base.cpu().numpy(), centers.cpu().numpy(), cid.to(torch.int32).cpu().numpy().astype(np.uint32))
    build_s = time.perf_counter() - tb
    log(f"build: {build_s:.2f} s ({args.n / build_s / 1e6:.2f} M vectors/s incl. host upload)")

    nq, k = args.nq, args.k
    q_dev = queries.contiguous()
    ids_dev = torch.empty((nq, k), dtype=torch.int32, device="cuda:0")
    d_dev = torch.empty((nq, k), dtype=torch.float32, device="cuda:0")
    blocks = ((args.dim + 63) // 64 * 64) // 8
    batch_bytes = (128 + 32 * blocks + 63) // 64 * 64
    survivor_bytes = blocks * (args.bits - 1) + 12
    results = {}
    for mode in modes:
        rows = []
        # "<form>_graph": the same form replayed as a HIP graph (the warmup
        # calls run it eagerly, then capture it).
        form = mode[:-len("_graph")] if mode.endswith("_graph") else mode
        e8qlib.set_search_graphs(mode != form)
        for nprobe in grid:
            for _ in range(max(args.warmup, 2 if mode != form else 0)):
                idx.search_device(q_dev.data_ptr(), nq, k, nprobe, ids_dev.data_ptr(), d_dev.data_ptr(), form)
            torch.cuda.synchronize()
            times, scan_ms, st = [], [], None
            for _ in range(args.steps):
                torch.cuda.synchronize()
                ts = time.perf_counter()
                st = idx.search_device(q_dev.data_ptr(), nq, k, nprobe, ids_dev.data_ptr(), d_dev.data_ptr(), form)
                torch.cuda.synchronize()
                times.append(time.perf_counter() - ts)
                scan_ms.append(st["device_scan_ms"])
            ids = ids_dev.cpu().numpy().view(np.uint32)
            ms = 1000.0 * sum(times) / len(times)
            row = {"nprobe": nprobe, "ms_per_step": ms, "qps": nq / (ms / 1000.0),
                   "recall": recall_at(gt, ids, 10, 10) if gt is not None else -1.0,
                   "recall_10at100": recall_at(gt, ids, 10, k) if gt is not None else -1.0,
                   "scan_kernel_ms": sum(scan_ms) / len(scan_ms), "batches": st["batches"],
                   "records_read": st["scanned"],
                   "survivors": st["survivors"], "candidates": st["candidates"]}
            algo = row["batches"] * batch_bytes + row["survivors"] * survivor_bytes
            row["scan_algorithmic_bytes"] = algo
            row["scan_achieved_gbs"] = algo / (row["scan_kernel_ms"] / 1000.0) / 1e9
            rows.append(row)
            log(f"{mode} nprobe={nprobe}: {ms:.2f} ms/step, {row['qps']:.0f} QPS, recall@10 (10@10)={row['recall']:.4f} "
                f"(10@{k} {row['recall_10at100']:.4f}), "
                f"scan kernel {row['scan_kernel_ms']:.2f} ms = {row['scan_achieved_gbs']:.0f} GB/s, "
                f"records read {row['records_read']}/{row['batches']}")
        results[mode] = {"rows": rows, "qps_at_target": qps_at_recall(rows, args.target) if gt is not None else None}
        if mode != form:
            results[mode]["graph"] = e8qlib.search_context_info()
            e8qlib.set_search_graphs(False)
    if args.profile:
        return
    best = max(results, key=lambda m: results[m]["qps_at_target"] or 0.0)
    rows = results[best]["rows"]
    # Operating point: the smallest grid nprobe reaching the target.
    op = next((r for r in rows if r["recall"] >= args.target), rows[-1])
    value = results[best]["qps_at_target"]

    # Single-query latency (the drop-in Search(query) path), plain form.
    lat = []
    qh = queries[:args.latency_queries].cpu().numpy()
    for i in range(qh.shape[0]):
        ts = time.perf_counter()
        idx.search(qh[i:i + 1], k, op["nprobe"])
        lat.append((time.perf_counter() - ts) * 1e6)
    lat.sort()
    p50 = lat[len(lat) // 2]
    p99 = lat[min(len(lat) - 1, int(0.99 * len(lat)))]

    cpu = None
    if not args.no_cpu:
        plain_ids = None
        idx.search_device(q_dev.data_ptr(), nq, k, op["nprobe"], ids_dev.data_ptr(), d_dev.data_ptr(), "plain")
        torch.cuda.synchronize()
        plain_ids = ids_dev.cpu().numpy().view(np.uint32)
        cpu = cpu_baseline(idx, queries.cpu().numpy(), k, op["nprobe"], plain_ids, args)
        log(f"cpu baseline: {cpu}")

    # Roofline of the plain scan kernel (its own events, eager launches) at
    # the largest grid nprobe (the scan-bound regime), and at the plain
    # form's operating point.
    rop = opp = op
    if "plain" in results:
        prow = results["plain"]["rows"]
        opp = next((r for r in prow if r["recall"] >= args.target), prow[-1])
        rop = max(prow, key=lambda r: r["nprobe"])
    achieved = rop["scan_achieved_gbs"]
    # HBM bytes per launch of the plain kernel at that nprobe, from the
    # committed rocprofv3 summary of this workload (copy-calibrated
    # FETCH_SIZE + WRITE_SIZE), when one covers it
    traffic = profile_traffic("PlainSearchKernel", f"e8q_{args.n // 1_000_000}Mx{args.dim}", rop["nprobe"])
    line = {
        "metric": "QPS at recall@10>=0.95 (plain vs batched/served, 1 GPU); p50/p99 latency",
        "value": round(value, 1) if value == value else None,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * nq / value, 3) if value == value and value > 0 else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 LUT / i32 coarse sums, fp32 refine",
        "data": "synthetic Gaussian mixture generated on the GPU (seed 42, centers N(0, 4I), points center + N(0, I))",
        "config": {"workload": "plain GPU E8Q scan: N=1M, D=128 Gaussian mixture, 1024 lists, b=4, "
                               f"10k queries, k=100, nprobe {args.grid}",
                   "n": args.n, "dim": args.dim, "nlist": args.nlist, "nq": nq, "k": k, "bits": args.bits,
                   "nprobe_grid": grid, "form": best, "operating_nprobe": op["nprobe"],
                   "recall_at_operating_point": op["recall"],
                   "recall_definition": "recall@10 = |exact top-10 (fp64) & first 10 results| / 10, the reference's "
                                        "recall_at_k (experiments/protocol.py:482-488); 10@100 per grid row as "
                                        "recall_10at100.  On this data 10@10 is capped near 0.91 by the 4-bit codes "
                                        "(the reference CPU search too), so value is null: the target is never "
                                        "reached"},
        "latency_us": {"p50": round(p50, 1), "p99": round(p99, 1), "what": "single-query search() calls"},
        "modes": {m: {"qps_at_target": results[m]["qps_at_target"],
                      "grid": [{kk: (round(v, 4) if isinstance(v, float) else v) for kk, v in r.items()}
                               for r in results[m]["rows"]]} for m in results},
        "build_seconds": round(build_s, 2),
        "roofline": {"bound": "hbm", "kernel": "PlainSearchKernel", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "basis": f"plain form at nprobe={rop['nprobe']}, per launch: {rop['batches']} scan records x "
                              f"{batch_bytes} B + {rop['survivors']} survivors x {survivor_bytes} B, / mean kernel "
                              f"time {rop['scan_kernel_ms']:.3f} ms",
                     "at_operating_point": {"nprobe": opp["nprobe"], "achieved": round(opp["scan_achieved_gbs"], 1),
                                            "frac": round(opp["scan_achieved_gbs"] / HBM_PEAK_GBS, 4)}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
