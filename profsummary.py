"""Turns rocprofv3 CSV output into the short summaries kept under profiles/.

    python profsummary.py stats  <kernel_stats.csv> <out.txt> <command...>
    python profsummary.py pmc    <fetch_counter_collection.csv>
                                                <write_counter_collection.csv>
                                                <kernel-substring> <copy-bytes> <out.json>
    python profsummary.py route  <fetch.csv> <write.csv> <copy-bytes> <transposes>
                                 <workload> <out.json> <command>...

    python profsummary.py counters <counter_collection.csv> <out.txt> <command...>

`counters` averages every collected counter per dispatch of each library kernel (and of the
calibration copy) and adds the LDS bank-conflict ratio
SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT) (stall cycles per cycle of
useful LDS work) and the stall cycles per LDS instruction, where those counters are present.

`pmc` scales the FETCH_SIZE / WRITE_SIZE readings (KiB per dispatch) of the named kernel by
the factor that maps the same counter on the calibration copy kernel (__amd_rocclr_copyBuffer,
a device copy of <copy-bytes> bytes: it reads and writes exactly that many) to its known
traffic, and reports HBM bytes per launch. `route` does the same for every library kernel of a
run of full transposes and reports HBM bytes per transpose.
"""
import csv
import json
import statistics
import sys
from collections import defaultdict


def counters(csv_path, out_path, command):
    rows = list(csv.DictReader(open(csv_path)))
    # (kernel, dispatch) -> counter -> value (a counter can come as several rows: sum them)
    per = defaultdict(lambda: defaultdict(float))
    for r in rows:
        name = r["Kernel_Name"]
        if "lm::" not in name and "copyBuffer" not in name:
            continue
        per[(name, r["Dispatch_Id"])][r["Counter_Name"]] += float(r["Counter_Value"])
    by_kernel = defaultdict(list)
    for (name, _), vals in per.items():
        by_kernel[name].append(vals)
    lines = ["# rocprofv3 counters, average per dispatch", f"# command: {command}", ""]
    for name, ds in sorted(by_kernel.items(), key=lambda kv: -len(kv[1])):
        keys = sorted({k for d in ds for k in d})
        avg = {k: statistics.mean(d.get(k, 0.0) for d in ds) for k in keys}
        lines.append(f"{name[:110]}  ({len(ds)} dispatches)")
        for k in keys:
            lines.append(f"    {k:32s} {avg[k]:16.1f}")
        bc, act = avg.get("SQ_LDS_BANK_CONFLICT"), avg.get("SQ_LDS_IDX_ACTIVE")
        if bc is not None and act and act > bc:
            lines.append(f"    {'bank-conflict stall / LDS work':32s} {bc / (act - bc):16.4f}")
        n = avg.get("SQ_INSTS_LDS")
        if bc is not None and n:
            lines.append(f"    {'stall cycles per LDS instruction':32s} {bc / n:16.3f}")
        lines.append("")
    open(out_path, "w").write("\n".join(lines))


def stats(csv_path, out_path, command):
    rows = list(csv.DictReader(open(csv_path)))
    lines = [f"# rocprofv3 kernel statistics", f"# command: {command}", "",
             f"{'avg_us':>10} {'min_us':>10} {'max_us':>10} {'calls':>6}  kernel"]
    for r in rows:
        lines.append(f"{float(r['AverageNs']) / 1e3:10.2f} {float(r['MinNs']) / 1e3:10.2f} "
                     f"{float(r['MaxNs']) / 1e3:10.2f} {int(r['Calls']):6d}  {r['Name'][:110]}")
    open(out_path, "w").write("\n".join(lines) + "\n")


def _per_kernel(path, counter):
    vals = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return vals


def pmc(fetch_csv, write_csv, kernel_sub, copy_bytes, out_path):
    f = _per_kernel(fetch_csv, "FETCH_SIZE")
    w = _per_kernel(write_csv, "WRITE_SIZE")

    def pick(d, sub):
        ks = [k for k in d if sub in k]
        if not ks:
            raise SystemExit(f"no kernel matching {sub!r}")
        return ks[0], statistics.mean(d[ks[0]])

    kname, kf = pick(f, kernel_sub)
    _, kw = pick(w, kernel_sub)
    _, cf = pick(f, "__amd_rocclr_copyBuffer")
    _, cw = pick(w, "__amd_rocclr_copyBuffer")
    fs = copy_bytes / (cf * 1024.0)
    ws = copy_bytes / (cw * 1024.0)
    res = {
        "kernel": kname,
        "fetch_kib_per_dispatch": kf, "write_kib_per_dispatch": kw,
        "copy_fetch_kib": cf, "copy_write_kib": cw, "copy_bytes": copy_bytes,
        "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
        "hbm_read_bytes_per_launch": round(kf * 1024 * fs),
        "hbm_write_bytes_per_launch": round(kw * 1024 * ws),
    }
    res["hbm_bytes_per_launch"] = res["hbm_read_bytes_per_launch"] + \
        res["hbm_write_bytes_per_launch"]
    json.dump(res, open(out_path, "w"), indent=1)
    return res


def route(fetch_csv, write_csv, copy_bytes, calls, workload, out_path, commands):
    """HBM traffic of every library kernel (names with "lm::") in a run of `calls` full
    transposes, per transpose, scaled by the calibration copy in the same run."""
    f = _per_kernel(fetch_csv, "FETCH_SIZE")
    w = _per_kernel(write_csv, "WRITE_SIZE")
    cf = statistics.mean(next(v for k, v in f.items() if "__amd_rocclr_copyBuffer" in k))
    cw = statistics.mean(next(v for k, v in w.items() if "__amd_rocclr_copyBuffer" in k))
    fs, ws = copy_bytes / (cf * 1024.0), copy_bytes / (cw * 1024.0)
    per = {}
    for k in f:
        if "lm::" not in k:
            continue
        rd = sum(f[k]) * 1024 * fs
        wr = sum(w.get(k, [0.0])) * 1024 * ws
        per[k[:90]] = {"dispatches": len(f[k]), "hbm_read_bytes": round(rd / calls),
                       "hbm_write_bytes": round(wr / calls)}
    tot = sum(v["hbm_read_bytes"] + v["hbm_write_bytes"] for v in per.values())
    res = {"workload": workload, "commands": commands, "transposes": calls,
           "calibration": f"one {copy_bytes}-byte __amd_rocclr_copyBuffer in the same run: "
                          f"FETCH_SIZE x{fs:.4f}, WRITE_SIZE x{ws:.4f}",
           "hbm_bytes_per_transpose": tot, "per_kernel_per_transpose": per}
    json.dump(res, open(out_path, "w"), indent=1)
    return res


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2], sys.argv[3], " ".join(sys.argv[4:]))
    elif sys.argv[1] == "pmc":
        print(pmc(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5]), sys.argv[6]))
    elif sys.argv[1] == "counters":
        counters(sys.argv[2], sys.argv[3], " ".join(sys.argv[4:]))
    elif sys.argv[1] == "route":
        # route <fetch.csv> <write.csv> <copy-bytes> <transposes> <workload> <out.json> <cmd>...
        print(route(sys.argv[2], sys.argv[3], int(sys.argv[4]), int(sys.argv[5]), sys.argv[6],
                    sys.argv[7], sys.argv[8:]))
