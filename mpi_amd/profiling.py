"""Summaries of rocprofv3 output for the kernels of libhz.so.

    python -m mpi_amd.profiling stats  <results.db | kernel_stats.csv>  [--top 12]
    python -m mpi_amd.profiling traffic --fetch <results.db | counter_collection.csv> \\
        --write <results.db | counter_collection.csv> --copy-mib M [--out summary.json]

rocprofv3 writes an SQLite database (`<name>_results.db`) by default and
CSV files with `--output-format csv`; both are read.

`stats` prints the per-kernel table of a `rocprofv3 --kernel-trace --stats`
run.  `traffic` reduces two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
runs of `bench.py --copy-probe-mib M` to HBM bytes per dispatch for every
kernel.  Each counter is scaled by what it reads for the probe, a plain
element-wise copy that reads M MiB and writes M MiB (the first
element-wise multiply in the run), so the bytes are in units the copy
defines.
"""
import argparse
import csv
import json
import sqlite3
import statistics
import sys


def _rows(path):
    with open(path, newline="") as fp:
        return list(csv.DictReader(fp))


def short_name(name):
    """Template-qualified kernel name without namespaces and argument list."""
    base = name.replace("(anonymous namespace)::", "")
    if base.startswith("void "):
        base = base[5:]
    return base.split("(")[0].replace("hz::", "")


def kernel_stats(path):
    out = []
    if path.endswith(".db"):
        db = sqlite3.connect(path)
        for name, calls, total, avg, pct in db.execute(
                "select name, total_calls, total_duration, average, percentage from top_kernels"):
            # the database's durations are in microseconds
            out.append({"kernel": short_name(name), "calls": int(calls), "avg_us": float(avg),
                        "total_ms": float(total) / 1e3, "percent": float(pct)})
        return out
    for r in _rows(path):
        out.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
                    "avg_us": float(r["AverageNs"]) / 1e3, "total_ms": float(r["TotalDurationNs"]) / 1e6,
                    "percent": float(r["Percentage"])})
    return out


def counter_by_kernel(path, counter):
    """{kernel: [value per dispatch]} for one counter (values in KiB for
    FETCH_SIZE / WRITE_SIZE)."""
    per = {}
    if path.endswith(".db"):
        db = sqlite3.connect(path)
        q = ("select k.name, p.counter_value from pmc_events p join kernels k on k.dispatch_id = p.dispatch_id "
             "where p.counter_name = ? order by p.dispatch_id")
        for name, val in db.execute(q, (counter,)):
            per.setdefault(name, []).append(float(val))
        return per
    for r in _rows(path):
        if r.get("Counter_Name") != counter:
            continue
        per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return per


def counters_summary(paths, prefix):
    """Median per-dispatch value of every counter in the given rocprofv3
    databases, for kernels whose short name starts with `prefix`, plus
    derived ratios when their inputs are present (MFMA busy share of the
    SIMD cycles, wait share of the wave cycles, LDS bank conflicts per LDS
    access cycle)."""
    out = {}
    for path in paths:
        db = sqlite3.connect(path)
        q = ("select k.name, p.counter_name, p.counter_value from pmc_events p join kernels k "
             "on k.dispatch_id = p.dispatch_id")
        vals = {}
        for name, cname, v in db.execute(q):
            sn = short_name(name)
            if sn.startswith(prefix):
                vals.setdefault(sn, {}).setdefault(cname, []).append(float(v))
        for sn, cs in vals.items():
            d = out.setdefault(sn, {})
            for cname, vs in cs.items():
                d[cname] = statistics.median(vs)
                d.setdefault("dispatches", len(vs))
    for d in out.values():
        if "SQ_VALU_MFMA_BUSY_CYCLES" in d and "SQ_BUSY_CU_CYCLES" in d and d["SQ_BUSY_CU_CYCLES"]:
            # busy-CU cycles are per CU; four SIMDs each carry one MFMA pipe
            d["mfma_busy_per_simd_cycle"] = d["SQ_VALU_MFMA_BUSY_CYCLES"] / (4 * d["SQ_BUSY_CU_CYCLES"])
        if "SQ_WAIT_ANY" in d and "SQ_WAVE_CYCLES" in d and d["SQ_WAVE_CYCLES"]:
            d["wait_any_per_wave_cycle"] = d["SQ_WAIT_ANY"] / d["SQ_WAVE_CYCLES"]
        if "SQ_WAIT_INST_LDS" in d and "SQ_WAVE_CYCLES" in d and d["SQ_WAVE_CYCLES"]:
            d["wait_lds_per_wave_cycle"] = d["SQ_WAIT_INST_LDS"] / d["SQ_WAVE_CYCLES"]
        if "SQ_LDS_BANK_CONFLICT" in d and "SQ_LDS_IDX_ACTIVE" in d and d["SQ_LDS_IDX_ACTIVE"]:
            d["lds_bank_conflict_per_active"] = d["SQ_LDS_BANK_CONFLICT"] / d["SQ_LDS_IDX_ACTIVE"]
    return out


def _probe_value(per):
    for name, vals in per.items():
        if "MulFunctor" in name:
            return vals[0]
    raise SystemExit("copy probe (element-wise multiply) not found in the counter file")


def traffic(fetch_csv, write_csv, copy_mib):
    fetch = counter_by_kernel(fetch_csv, "FETCH_SIZE")
    write = counter_by_kernel(write_csv, "WRITE_SIZE")
    kib = copy_mib * 1024.0
    f_scale = kib / _probe_value(fetch)
    w_scale = kib / _probe_value(write)
    kernels = {}
    for name in sorted(set(fetch) | set(write)):
        if not name.startswith("void hz::"):
            continue
        fv = fetch.get(name, [])
        wv = write.get(name, [])
        kernels[short_name(name)] = {
            "dispatches": max(len(fv), len(wv)),
            "fetch_kib_median": statistics.median(fv) if fv else None,
            "write_kib_median": statistics.median(wv) if wv else None,
            "hbm_bytes_per_dispatch": (statistics.median(fv) * f_scale + statistics.median(wv) * w_scale) * 1024.0
            if fv and wv else None,
        }
    return {"copy_probe_mib": copy_mib, "fetch_scale": f_scale, "write_scale": w_scale, "kernels": kernels}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=12)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--out")
    c = sub.add_parser("counters")
    c.add_argument("db", nargs="+")
    c.add_argument("--kernel", default="", help="short-name prefix of the kernels to report")
    a = ap.parse_args(argv)
    if a.cmd == "counters":
        print(json.dumps(counters_summary(a.db, a.kernel), indent=1, sort_keys=True))
        return
    if a.cmd == "stats":
        rows = kernel_stats(a.csv)[: a.top]
        print(f"{'kernel':70s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'%':>6s}")
        for r in rows:
            print(f"{r['kernel'][:70]:70s} {r['calls']:6d} {r['avg_us']:10.2f} {r['total_ms']:10.3f} {r['percent']:6.2f}")
    else:
        res = traffic(a.fetch, a.write, a.copy_mib)
        txt = json.dumps(res, indent=1)
        if a.out:
            with open(a.out, "w") as fp:
                fp.write(txt + "\n")
        print(txt)
    return 0


if __name__ == "__main__":
    sys.exit(main())
