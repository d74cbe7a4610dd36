"""Writes the golden fixtures under tests/golden/ from the reference library.

Run in the build container only (needs /root/reference and the reference
build made by `make -f oracle/ref.mk`):

    python tests/golden/make_golden.py

It drives oracle/_ref/gen_golden (tests/golden/gen_golden.cpp linked against
the reference e8qlib) on the seeded inputs of cases.py and keeps its outputs
as .npy/.e8ax fixtures plus manifest.json.  `--only base,flush,c1,c3,diag` writes
only the named sections (default: all).
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import cases  # noqa: E402

GEN = os.path.join(HERE, "..", "..", "oracle", "_ref", "gen_golden")


def run(*args, env=None):
    subprocess.run([GEN] + [str(a) for a in args], check=True,
                   env=dict(os.environ, **env) if env else None)


def base_fixtures(out, manifest, t):
    """Level tables, rotations, encoder digits, coarse sums, route lists,
    the E8AX and ME8QIDXX fixture indexes and their searches."""
    manifest.update({"index": {}, "rotate": [], "inputs": {}})
    # Level tables, packed per rate: for b = 1..13, branch 0 then branch 1.
    run("levels", t("lev.npy"), t("scale.npy"), t("wcls.npy"))
    lev = np.load(t("lev.npy"))
    flat = np.concatenate([lev[b - 1, br, : 1 << b] for b in range(1, 14) for br in (0, 1)])
    np.save(os.path.join(out, "levels.npy"), flat)
    np.save(os.path.join(out, "fixed_scale.npy"), np.load(t("scale.npy")))
    np.save(os.path.join(out, "class_means.npy"), np.load(t("wcls.npy")))
    for dim, seed in cases.ROTATE_CASES:
        x = cases.rotate_input(dim)
        np.save(t("x.npy"), x)
        run("rotate", dim, seed, t("x.npy"), t("s.npy"), t("r.npy"))
        np.savez(os.path.join(out, f"rotate_{dim}.npz"), signs=np.load(t("s.npy")), out=np.load(t("r.npy")))
        manifest["rotate"].append([dim, seed])
    u = cases.encode_input()
    np.save(t("u.npy"), u)
    enc = {}
    for b in range(1, 14):
        run("encode", b, cases.ENCODE_PADDED, t("u.npy"), t("d.npy"), t("b.npy"))
        enc[f"digits_b{b}"] = np.load(t("d.npy"))
        enc[f"branch_b{b}"] = np.load(t("b.npy"))
    np.savez(os.path.join(out, "encode.npz"), **enc)
    manifest["inputs"]["encode"] = cases.digest(u)
    acc = {}
    for blocks in cases.ACCUM_BLOCKS:
        pay, lut = cases.accum_input(blocks)
        np.save(t("p.npy"), pay)
        np.save(t("l.npy"), lut)
        run("accum", t("p.npy"), t("l.npy"), t("o1.npy"), t("o2.npy"))
        acc[f"single_{blocks}"] = np.load(t("o1.npy"))
        acc[f"pair_{blocks}"] = np.load(t("o2.npy"))
    np.savez(os.path.join(out, "accum.npz"), **acc)
    q, c = cases.route_input()
    np.save(t("q.npy"), q)
    np.save(t("c.npy"), c)
    run("route", t("q.npy"), t("c.npy"), 20, t("rd.npy"), t("ri.npy"), t("rn.npy"))
    np.savez(os.path.join(out, "route.npz"), dist=np.load(t("rd.npy")), ids=np.load(t("ri.npy")),
             cn2=np.load(t("rn.npy")))
    for name, (n, dim, nlist, bits, seed, alpha, _e, _nq) in cases.INDEX_CASES.items():
        data, cent, lab, queries = cases.mixture(name)
        for fn, arr in (("data", data), ("cent", cent), ("lab", lab), ("q", queries)):
            np.save(t(fn + ".npy"), arr)
        idx_path = os.path.join(out, f"{name}.e8ax")
        run("build", t("data.npy"), t("cent.npy"), t("lab.npy"), bits, seed, alpha, idx_path)
        res = {}
        for (k, nprobe, dis) in cases.search_grid(name):
            pre = t("s")
            run("search", idx_path, t("q.npy"), k, nprobe, dis, pre)
            tag = f"k{k}_p{nprobe}_d{dis}"
            for part in ("ids", "dists", "stats", "rotq", "lut", "consts", "probe", "pdist"):
                res[f"{tag}_{part}"] = np.load(f"{pre}_{part}.npy")
        res["queries"] = queries
        np.savez(os.path.join(out, f"{name}_search.npz"), **res)
        manifest["index"][name] = cases.digest(data, cent, lab)
    manifest["me8q"] = {}
    for name, (n, dim, nlist, avg, seed, alpha, spec, pv, de, _e, _nq) in cases.ME8Q_CASES.items():
        data, cent, lab, queries, lam = cases.me8q_mixture(name)
        for fn, arr in (("data", data), ("cent", cent), ("lab", lab), ("var", lam), ("mq", queries)):
            np.save(t(fn + ".npy"), arr)
        idx_path = os.path.join(out, f"{name}.me8q")
        pre = t("m")
        run("mbuild", t("data.npy"), t("cent.npy"), t("lab.npy"), avg, seed, alpha, spec or "-",
            t("var.npy") if pv else "-", int(de), idx_path, pre)
        res = {part: np.load(f"{pre}_{part}.npy")
               for part in ("plane", "refine", "gain", "vpc", "rcn2", "off", "cn2", "crot", "plan", "routedim")}
        for (k, nprobe) in cases.me8q_search_grid(name):
            run("msearch", idx_path, t("mq.npy"), k, nprobe, pre)
            tag = f"k{k}_p{nprobe}"
            for part in ("ids", "dists", "stats", "rotq", "probe", "pdist"):
                res[f"{tag}_{part}"] = np.load(f"{pre}_{part}.npy")
        res["queries"] = queries
        np.savez_compressed(os.path.join(out, f"{name}_me8q.npz"), **res)
        manifest["me8q"][name] = cases.digest(data, cent, lab, lam)


def flush_fixtures(out, manifest, t):
    """Searches of the E8Q fixture indexes under E8Q_FLUSH_EVERY = each of
    cases.FLUSH_SIZES (the reference reads it once per process)."""
    res = {}
    for name in cases.FLUSH_INDEXES:
        _d, _c, _l, queries = cases.mixture(name)
        np.save(t("q.npy"), queries)
        for fe in cases.FLUSH_SIZES:
            for (k, nprobe, dis) in cases.flush_grid(name):
                pre = t("s")
                run("search", os.path.join(out, f"{name}.e8ax"), t("q.npy"), k, nprobe, dis, pre,
                    env={"E8Q_FLUSH_EVERY": str(fe)})
                tag = f"{name}_f{fe}_k{k}_p{nprobe}_d{dis}"
                for part in ("ids", "dists", "stats"):
                    res[f"{tag}_{part}"] = np.load(f"{pre}_{part}.npy")
    np.savez_compressed(os.path.join(out, "flush.npz"), **res)
    manifest["flush"] = {"sizes": cases.FLUSH_SIZES, "indexes": cases.FLUSH_INDEXES}


def c1_fixtures(out, manifest, t):
    """The 100k x 128 i.i.d. configuration: labels, per-section image digests
    of the reference's E8AX file, and the search of the 1000 queries."""
    sys.path.insert(0, os.path.dirname(HERE))
    import e8ax
    base, cent, queries = cases.c1_iid()
    lab = cases.nearest_centroid(base, cent).astype(np.uint32)
    for fn, arr in (("data", base), ("cent", cent), ("lab", lab), ("q", queries)):
        np.save(t(fn + ".npy"), arr)
    path = t("c1.e8ax")
    c = cases.C1
    run("build", t("data.npy"), t("cent.npy"), t("lab.npy"), c["bits"], c["seed"], 0.0, path)
    dig = cases.e8ax_digests(e8ax.read(path))
    pre = t("s")
    run("search", path, t("q.npy"), c["k"], c["nprobe"], 0, pre)
    res = {part: np.load(f"{pre}_{part}.npy") for part in ("ids", "dists", "stats", "probe")}
    res["labels"] = lab.astype(np.uint8)
    np.savez_compressed(os.path.join(out, "c1_iid.npz"), **res)
    manifest["c1"] = {"inputs": cases.digest(base, cent, queries), "image": dig}


def c3_fixtures(out, manifest, t):
    """The D = 960 ME8Q shape: labels, the reference's plan, the SHA-256 of
    its ME8QIDXX image, and searches over cases.c3_grid()."""
    data, cent, queries, lam = cases.c3_shape()
    lab = cases.nearest_centroid(data, cent).astype(np.uint32)
    for fn, arr in (("data", data), ("cent", cent), ("lab", lab), ("var", lam), ("mq", queries)):
        np.save(t(fn + ".npy"), arr)
    path = t("c3.me8q")
    pre = t("m")
    c = cases.C3
    run("mbuild", t("data.npy"), t("cent.npy"), t("lab.npy"), c["avg_bits"], c["seed"], 0.0, "-", t("var.npy"), 1,
        path, pre)
    res = {"plan": np.load(f"{pre}_plan.npy"), "labels": lab.astype(np.uint8)}
    for (k, nprobe) in cases.c3_grid():
        run("msearch", path, t("mq.npy"), k, nprobe, pre)
        for part in ("ids", "dists", "stats", "probe"):
            res[f"k{k}_p{nprobe}_{part}"] = np.load(f"{pre}_{part}.npy")
    np.savez_compressed(os.path.join(out, "c3_shape.npz"), **res)
    manifest["c3"] = {"inputs": cases.digest(data, cent, queries, lam),
                      "image_sha256": cases.digest(np.fromfile(path, np.uint8))}


def diag_fixtures(out, manifest, t):
    """The diagnostics over the fixture indexes: E8Q Estimate, ProbeBoundW
    and ExhaustiveTopK; ME8Q Estimate, ExhaustiveTopK and SearchIVF."""
    res = {}
    for name in cases.INDEX_CASES:
        data, _c, _l, queries = cases.mixture(name)
        pq, pid = cases.diag_pairs(name)
        for fn, arr in (("q", queries), ("base", data), ("pq", pq), ("pid", pid)):
            np.save(t(fn + ".npy"), arr)
        for k in cases.diag_k(name):
            pre = t("g")
            run("diag", os.path.join(out, f"{name}.e8ax"), t("q.npy"), t("base.npy"), t("pq.npy"), t("pid.npy"), k,
                pre)
            res[f"{name}_est"] = np.load(f"{pre}_est.npy")
            res[f"{name}_pbw"] = np.load(f"{pre}_pbw.npy")
            res[f"{name}_k{k}_exids"] = np.load(f"{pre}_exids.npy")
            res[f"{name}_k{k}_exd"] = np.load(f"{pre}_exd.npy")
    for name in cases.ME8Q_CASES:
        _d, _c, _l, queries, _lam = cases.me8q_mixture(name)
        pq, pid = cases.diag_pairs(name)
        for fn, arr in (("q", queries), ("pq", pq), ("pid", pid)):
            np.save(t(fn + ".npy"), arr)
        for (k, nprobe) in cases.diag_ivf_grid(name):
            pre = t("g")
            run("mdiag", os.path.join(out, f"{name}.me8q"), t("q.npy"), t("pq.npy"), t("pid.npy"), k, nprobe, pre)
            res[f"{name}_est"] = np.load(f"{pre}_est.npy")
            res[f"{name}_k{k}_exids"] = np.load(f"{pre}_exids.npy")
            res[f"{name}_k{k}_exd"] = np.load(f"{pre}_exd.npy")
            for part in ("ivfids", "ivfd", "ivfstats"):
                res[f"{name}_k{k}_p{nprobe}_{part}"] = np.load(f"{pre}_{part}.npy")
    np.savez_compressed(os.path.join(out, "diag.npz"), **res)
    manifest["diag"] = {"e8q": list(cases.INDEX_CASES), "me8q": list(cases.ME8Q_CASES)}


SECTIONS = {"base": base_fixtures, "flush": flush_fixtures, "c1": c1_fixtures, "c3": c3_fixtures,
            "diag": diag_fixtures}


def main():
    if not os.path.exists(GEN):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    only = sys.argv[sys.argv.index("--only") + 1].split(",") if "--only" in sys.argv else list(SECTIONS)
    out = HERE
    mpath = os.path.join(out, "manifest.json")
    manifest = json.load(open(mpath)) if os.path.exists(mpath) else {}
    with tempfile.TemporaryDirectory() as tmp:
        t = lambda n: os.path.join(tmp, n)  # noqa: E731
        for name in only:
            SECTIONS[name](out, manifest, t)
    with open(mpath, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
