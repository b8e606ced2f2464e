"""Summaries of rocprofv3 result databases (rocpd SQLite) for the profiles/ folder.

    python -m phg.profile_summary stats  <kernel_trace.db> [--skip K]
    python -m phg.profile_summary pmc    <fetch.db> <write.db> <copy_fetch.db> <copy_write.db> <copy_bytes>
    python -m phg.profile_summary l2     <tcc.db>

    python -m phg.profile_summary csvstats <kernel_trace.csv> [--skip K]
    python -m phg.profile_summary counters <counter_collection.csv> [--skip K]
    python -m phg.profile_summary profile <root> <leaf,width,block,paths> ...

`csvstats` / `counters` read rocprofv3's CSV output (--output-format csv): per-kernel
durations from the kernel trace, and per-kernel averages over dispatches of every collected
counter (rows of one dispatch and counter, one per hardware instance, summed first).
`stats` prints per-kernel call counts and durations (with --skip K, leaving out each kernel's
first K dispatches, e.g. those of a counted warm-up render); `pmc` prints per-kernel FETCH_SIZE and
WRITE_SIZE per dispatch, scaled by what the same counters read for a plain device copy of
`copy_bytes` bytes, and emits a JSON line with the HBM traffic per launch of each kernel;
`l2` prints each kernel's L2 hit rate from TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum).
"""
import json
import re
import sqlite3
import sys


def _rows(path, sql, args=()):
    con = sqlite3.connect(path)
    try:
        return con.execute(sql, args).fetchall()
    finally:
        con.close()


def short(name):
    return re.sub(r"\(.*", "", name)


def kernel_stats(path, skip=0):
    if skip <= 0:
        rows = _rows(path, "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                           "from kernels group by name order by sum(duration) desc")
    else:
        per = {}
        try:
            ordered = _rows(path, "select name, duration from kernels order by start")
        except sqlite3.OperationalError:  # no start column in this schema: dispatch (row) order
            ordered = _rows(path, "select name, duration from kernels order by rowid")
        for name, d in ordered:
            per.setdefault(name, []).append(d)
        rows = []
        for name, ds in per.items():
            ds = ds[skip:]
            if ds:
                rows.append((name, len(ds), sum(ds), sum(ds) / len(ds), min(ds), max(ds)))
        rows.sort(key=lambda r: -r[2])
    total = sum(r[2] for r in rows) or 1
    return [{"kernel": short(n), "calls": c, "total_ms": s / 1e6, "avg_ms": a / 1e6, "min_ms": lo / 1e6,
             "max_ms": hi / 1e6, "percent": 100.0 * s / total} for n, c, s, a, lo, hi in rows]


def counter(path, name):
    rows = _rows(path, "select name, count(*), avg(counter_value) from pmc_events where counter_name=? group by name",
                 (name,))
    return {short(n): (c, a) for n, c, a in rows}


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def csv_kernel_stats(path, skip=0):
    per = {}
    rows = _csv_rows(path)
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    for r in rows:
        per.setdefault(short(r["Kernel_Name"]), []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    out = []
    for k, ds in per.items():
        ds = ds[skip:]
        if ds:
            out.append({"kernel": k, "calls": len(ds), "total_ms": sum(ds), "avg_ms": sum(ds) / len(ds),
                        "min_ms": min(ds), "max_ms": max(ds)})
    out.sort(key=lambda r: -r["total_ms"])
    total = sum(r["total_ms"] for r in out) or 1.0
    for r in out:
        r["percent"] = 100.0 * r["total_ms"] / total
    return out


def csv_counters(path, skip=0):
    """{kernel: {"dispatches": n, counter: average per dispatch}}"""
    per_dispatch = {}
    order = {}
    for r in _csv_rows(path):
        d = int(r["Dispatch_Id"])
        k = short(r["Kernel_Name"])
        order[d] = k
        key = (d, r["Counter_Name"])
        per_dispatch[key] = per_dispatch.get(key, 0.0) + float(r["Counter_Value"])
    seen, keep = {}, set()
    for d in sorted(order):
        k = order[d]
        seen[k] = seen.get(k, 0) + 1
        if seen[k] > skip:
            keep.add(d)
    out = {}
    for (d, c), v in per_dispatch.items():
        if d not in keep:
            continue
        e = out.setdefault(order[d], {"dispatches": set()})
        e["dispatches"].add(d)
        e.setdefault(c, []).append(v)
    res = {}
    for k, e in out.items():
        res[k] = {"dispatches": len(e["dispatches"])}
        for c, vs in e.items():
            if c != "dispatches":
                res[k][c] = sum(vs) / len(vs)
    return res


def counter_profile(root, configs, skip=9, cus=256, xcds=8, clock_ghz=2.4, copy_bytes=1 << 30):
    """Per-kernel, per-launch figures of the headline workload for each launch configuration
    (directories <root>/<leaf_width_block_paths>/{kt,valu,fetch,write,l2} of rocprofv3 CSV
    output, and <root>/{cf,cw} for the copy calibration), skipping each kernel's first `skip`
    launches (the counted 1-spp render): trace time, VALU and SALU instructions, VALU busy
    fraction (SQ_ACTIVE_INST_VALU / CUs / GUI-active cycles of one XCD), VALU lane utilization
    (SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)), bytes past L2 (copy-calibrated
    FETCH_SIZE and WRITE_SIZE) and L2 hit rate."""
    import glob
    import os

    def one(path_glob):
        hits = glob.glob(path_glob)
        return hits[0] if hits else None
    cf = csv_counters(one(f"{root}/cf/*counter_collection.csv"))
    cw = csv_counters(one(f"{root}/cw/*counter_collection.csv"))
    kib = copy_bytes / 1024.0
    kf = kib / max(v["FETCH_SIZE"] for v in cf.values())
    kw = kib / max(v["WRITE_SIZE"] for v in cw.values())
    out = {"compute_units": cus, "xcds": xcds, "peak_engine_clock_ghz": clock_ghz, "fetch_scale": kf,
           "write_scale": kw, "copy_bytes": copy_bytes, "skip_launches": skip, "configs": {}}
    for cfg in configs:
        d = os.path.join(root, cfg.replace(",", "_"))
        kt = {r["kernel"]: r for r in csv_kernel_stats(one(f"{d}/kt/*kernel_trace.csv"), skip)}
        valu = csv_counters(one(f"{d}/valu/*counter_collection.csv"), skip)
        f = csv_counters(one(f"{d}/fetch/*counter_collection.csv"), skip)
        w = csv_counters(one(f"{d}/write/*counter_collection.csv"), skip)
        l2 = csv_counters(one(f"{d}/l2/*counter_collection.csv"), skip)
        ks = {}
        for k, v in valu.items():
            if "phg::" not in k:
                continue
            gui = v["GRBM_GUI_ACTIVE"] / xcds
            e = {"launches": v["dispatches"], "insts_valu": v["SQ_INSTS_VALU"], "insts_salu": v["SQ_INSTS_SALU"],
                 "valu_busy": v["SQ_ACTIVE_INST_VALU"] / cus / gui,
                 "lane_utilization": v["SQ_THREAD_CYCLES_VALU"] / (64.0 * v["SQ_ACTIVE_INST_VALU"])}
            if k in kt:
                e["trace_avg_ms"] = kt[k]["avg_ms"]
            if k in f:
                e["fetch_bytes"] = f[k]["FETCH_SIZE"] * kf * 1024
                e["write_bytes"] = w.get(k, {}).get("WRITE_SIZE", 0.0) * kw * 1024
                e["traffic_bytes"] = e["fetch_bytes"] + e["write_bytes"]
            if k in l2:
                h, m = l2[k].get("TCC_HIT_sum", 0.0), l2[k].get("TCC_MISS_sum", 0.0)
                e["l2_hit_rate"] = h / (h + m) if h + m > 0 else None
            ks[k] = e
        out["configs"][cfg] = {"kernels": ks}
    return out


def main(argv):
    if len(argv) >= 3 and argv[0] == "profile":
        print(json.dumps(counter_profile(argv[1], argv[2:]), indent=1))
        return 0
    if len(argv) >= 2 and argv[0] == "csvstats":
        skip = int(argv[argv.index("--skip") + 1]) if "--skip" in argv else 0
        print(f"{'kernel':60s} {'calls':>6s} {'total ms':>10s} {'avg ms':>9s} {'min ms':>9s} {'max ms':>9s} {'%':>6s}")
        for r in csv_kernel_stats(argv[1], skip):
            print(f"{r['kernel'][:60]:60s} {r['calls']:6d} {r['total_ms']:10.3f} {r['avg_ms']:9.4f} "
                  f"{r['min_ms']:9.4f} {r['max_ms']:9.4f} {r['percent']:6.2f}")
        return 0
    if len(argv) >= 2 and argv[0] == "counters":
        skip = int(argv[argv.index("--skip") + 1]) if "--skip" in argv else 0
        res = csv_counters(argv[1], skip)
        names = sorted({c for e in res.values() for c in e if c != "dispatches"})
        print(f"{'kernel':36s} {'disp':>5s} " + " ".join(f"{n[:22]:>22s}" for n in names))
        for k in sorted(res, key=lambda k: -res[k].get(names[0], 0) if names else 0):
            e = res[k]
            print(f"{k[:36]:36s} {e['dispatches']:5d} " + " ".join(f"{e.get(n, float('nan')):22.6g}" for n in names))
        print(json.dumps(res))
        return 0
    if len(argv) >= 2 and argv[0] == "stats":
        skip = int(argv[argv.index("--skip") + 1]) if "--skip" in argv else 0
        st = kernel_stats(argv[1], skip)
        print(f"{'kernel':60s} {'calls':>6s} {'total ms':>10s} {'avg ms':>9s} {'min ms':>9s} {'max ms':>9s} {'%':>6s}")
        for r in st:
            print(f"{r['kernel'][:60]:60s} {r['calls']:6d} {r['total_ms']:10.3f} {r['avg_ms']:9.4f} "
                  f"{r['min_ms']:9.4f} {r['max_ms']:9.4f} {r['percent']:6.2f}")
        return 0
    if len(argv) == 6 and argv[0] == "pmc":
        f, w, cf, cw = (counter(p, c) for p, c in zip(argv[1:5], ["FETCH_SIZE", "WRITE_SIZE", "FETCH_SIZE", "WRITE_SIZE"]))
        copy_kib = float(argv[5]) / 1024.0
        kf = copy_kib / max(v[1] for v in cf.values())
        kw = copy_kib / max(v[1] for v in cw.values())
        print(f"copy calibration: FETCH_SIZE x {kf:.4f}, WRITE_SIZE x {kw:.4f} (copy of {copy_kib:.0f} KiB)")
        out = {"fetch_scale": kf, "write_scale": kw, "kernels": {}}
        print(f"{'kernel':40s} {'dispatches':>10s} {'fetch MB':>10s} {'write MB':>10s} {'HBM MB/launch':>14s}")
        for k in sorted(f, key=lambda k: -f[k][1]):
            n, fa = f[k]
            wa = w.get(k, (0, 0.0))[1]
            fb, wb = fa * kf * 1024, wa * kw * 1024
            out["kernels"][k] = {"dispatches": n, "fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb}
            print(f"{k[:40]:40s} {n:10d} {fb / 1e6:10.2f} {wb / 1e6:10.2f} {(fb + wb) / 1e6:14.2f}")
        print(json.dumps(out))
        return 0
    if len(argv) == 2 and argv[0] == "l2":
        hit, miss = counter(argv[1], "TCC_HIT_sum"), counter(argv[1], "TCC_MISS_sum")
        print(f"{'kernel':40s} {'dispatches':>10s} {'hits/launch':>12s} {'misses/launch':>14s} {'L2 hit rate':>12s}")
        for k in sorted(hit, key=lambda k: -(hit[k][1] + miss.get(k, (0, 0.0))[1])):
            n, h = hit[k]
            m = miss.get(k, (0, 0.0))[1]
            rate = h / (h + m) if h + m > 0 else float("nan")
            print(f"{k[:40]:40s} {n:10d} {h:12.0f} {m:14.0f} {rate:12.4f}")
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
