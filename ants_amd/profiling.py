"""Summaries of rocprofv3 result databases for the SyN kernels.

``python -m ants_amd.profiling <results.db> [--copy-bytes B]`` prints, per
kernel and launch shape (grid size identifies the pyramid level), the call
count, total and mean duration, and -- when the run collected FETCH_SIZE /
WRITE_SIZE counters -- the mean HBM KiB per dispatch, scaled by what the same
counter reads for the calibration copy kernel of known size (bench.py
--copy-calibration launches one copying B bytes).

``python -m ants_amd.profiling traffic --fetch DIR --write DIR --kernel K
--copy-bytes B`` combines two counter passes written with ``-f csv`` (one
with FETCH_SIZE, one with WRITE_SIZE; the hardware cannot collect both in
one pass) into the HBM bytes per dispatch of kernel K, each counter scaled by
the ratio B / (its reading for the copy kernel), where the copy kernel reads
B bytes and writes B bytes.
"""
from __future__ import annotations

import argparse
import collections
import glob
import os
import sqlite3
import sys


def _short(name: str) -> str:
    base = name.replace("(anonymous namespace)::", "").replace("void ", "")
    base = base.split("(")[0]
    return base.replace("ants::", "")


def kernel_rows(db: str):
    c = sqlite3.connect(db)
    rows = c.execute("select id, name, grid_x, grid_y, grid_z, duration from kernels").fetchall()
    pmc = collections.defaultdict(dict)
    try:
        # counter runs: one row per (dispatch, counter); kernels.id follows dispatch order
        disp = {r[0]: r[1] for r in c.execute("select dispatch_id, id from kernels").fetchall()}
        for did, cname, val in c.execute("select dispatch_id, counter_name, value from counters_collection"):
            kid = disp.get(did)
            if kid is not None:
                pmc[kid][cname] = pmc[kid].get(cname, 0.0) + float(val)
    except sqlite3.Error:
        pass
    return rows, pmc


def summarize(db: str, copy_bytes: float = 0.0) -> str:
    rows, pmc = kernel_rows(db)
    agg = collections.OrderedDict()
    for kid, name, gx, gy, gz, dur in rows:
        key = (_short(name), gx, gy, gz)
        a = agg.setdefault(key, {"n": 0, "ns": 0.0, "ctr": collections.defaultdict(float)})
        a["n"] += 1
        a["ns"] += dur
        for cname, v in pmc.get(kid, {}).items():
            a["ctr"][cname] += v
    # calibration: the copy kernel moves copy_bytes in and out
    scale = {}
    if copy_bytes > 0:
        for (k, gx, gy, gz), a in agg.items():
            if k.startswith("k_copy"):
                for cname in ("FETCH_SIZE", "WRITE_SIZE"):
                    if a["ctr"].get(cname):
                        per = a["ctr"][cname] / a["n"] * 1024.0
                        scale[cname] = copy_bytes / per
    total = sum(a["ns"] for a in agg.values())
    lines = [f"{'kernel':<24} {'grid(x,y,z)':<22} {'calls':>6} {'total ms':>10} {'mean us':>10} {'%':>6}  counters/dispatch"]
    for (k, gx, gy, gz), a in sorted(agg.items(), key=lambda kv: -kv[1]["ns"]):
        extra = ""
        for cname in ("FETCH_SIZE", "WRITE_SIZE"):
            if cname in a["ctr"]:
                kib = a["ctr"][cname] / a["n"]
                mb = kib * 1024.0 * scale.get(cname, 1.0) / 1e6
                extra += f" {cname}={kib:.0f}KiB({mb:.1f}MB{' scaled' if cname in scale else ''})"
        lines.append(f"{k:<24} {str((gx, gy, gz)):<22} {a['n']:>6} {a['ns'] / 1e6:>10.3f} {a['ns'] / a['n'] / 1e3:>10.2f} "
                     f"{100.0 * a['ns'] / max(total, 1):>6.1f} {extra}")
    lines.append(f"total kernel time {total / 1e6:.3f} ms")
    if scale:
        lines.append("counter scale (copy kernel bytes / counter bytes): " +
                     ", ".join(f"{k}={v:.3f}" for k, v in scale.items()))
    return "\n".join(lines)


def counter_csv(path: str):
    """Rows (dispatch, kernel short name, grid, counter, value) of a rocprofv3
    counter_collection.csv (a file, or the newest one under a directory)."""
    import csv
    if os.path.isdir(path):
        found = sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True))
        if not found:
            raise FileNotFoundError(f"no counter_collection.csv under {path}")
        path = found[-1]
    out = []
    with open(path, newline="") as fh:
        rd = csv.DictReader(fh)
        cols = {c.lower(): c for c in rd.fieldnames or []}

        def col(*names):
            for n in names:
                if n in cols:
                    return cols[n]
            raise KeyError(f"{path}: none of {names} in {list(cols)}")

        c_id, c_name = col("dispatch_id"), col("kernel_name")
        c_grid, c_ctr, c_val = col("grid_size"), col("counter_name"), col("counter_value")
        for r in rd:
            out.append((int(r[c_id]), _short(r[c_name]), r[c_grid], r[c_ctr], float(r[c_val])))
    return out, path


def dispatch_grids(path: str) -> dict:
    """dispatch id -> "gx x gy x gz" from the kernel_trace.csv of the same
    rocprofv3 run (empty when the run traced no kernels).  The counter file
    only holds the total work-item count, which two pyramid levels can share."""
    import csv
    d = path if os.path.isdir(path) else os.path.dirname(path)
    found = sorted(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True))
    if not found:
        return {}
    out = {}
    with open(found[-1], newline="") as fh:
        rd = csv.DictReader(fh)
        cols = {c.lower(): c for c in rd.fieldnames or []}
        try:
            cid = cols["dispatch_id"]
            gx, gy, gz = cols["grid_size_x"], cols["grid_size_y"], cols["grid_size_z"]
        except KeyError:
            return {}
        for r in rd:
            out[int(r[cid])] = f"{r[gx]}x{r[gy]}x{r[gz]}"
    return out


def traffic(fetch_dir: str, write_dir: str, kernel: str, copy_bytes: float) -> str:
    lines = []
    per = {}
    for cname, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows, path = counter_csv(d)
        grids = dispatch_grids(path)
        lines.append(f"{cname}: {path}" + (" (grids from the run's kernel trace)" if grids else ""))
        kib = collections.defaultdict(float)  # (kernel, grid) -> summed KiB over dispatches
        disp = collections.defaultdict(set)
        for did, k, grid, ctr, v in rows:
            grid = grids.get(did, grid)
            if ctr == cname:
                kib[(k, grid)] += v
                disp[(k, grid)].add(did)
        copy = [(kib[key], len(disp[key])) for key in kib if key[0].startswith("k_copy_calibration")]
        if not copy:
            raise RuntimeError(f"{cname}: the pass holds no k_copy_calibration dispatch")
        copy_kib = sum(c for c, _ in copy) / sum(n for _, n in copy)
        scale = copy_bytes / (copy_kib * 1024.0)
        lines.append(f"  copy kernel: {copy_kib:.0f} KiB/dispatch for {copy_bytes / 2**20:.0f} MiB moved; "
                     f"scale {scale:.4f}")
        tot, nd = 0.0, 0
        for (k, grid), v in sorted(kib.items(), key=lambda kv: -kv[1]):
            if k.split("<")[0] == kernel:
                n = len(disp[(k, grid)])
                lines.append(f"  {k} grid {grid}: {n} dispatches, {v / n:.0f} KiB/dispatch "
                             f"({v / n * 1024 * scale / 1e6:.2f} MB scaled)")
                tot += v
                nd += n
        if nd == 0:
            raise RuntimeError(f"{cname}: no dispatch of {kernel}")
        per[cname] = tot / nd * 1024.0 * scale
        lines.append(f"  {kernel}: {nd} dispatches, mean {per[cname] / 1e6:.3f} MB/dispatch scaled")
    lines.append(f"kernel: {kernel}")
    lines.append(f"fetch_bytes_per_launch: {per['FETCH_SIZE']:.0f}")
    lines.append(f"write_bytes_per_launch: {per['WRITE_SIZE']:.0f}")
    lines.append(f"traffic_bytes_per_launch: {per['FETCH_SIZE'] + per['WRITE_SIZE']:.0f}")
    return "\n".join(lines)


def _trace_rows(path: str):
    """(dispatch id, short kernel name with template arguments, grid "x*y*z",
    duration ns) from the newest kernel_trace.csv under path, in dispatch order."""
    import csv
    if os.path.isdir(path):
        found = sorted(glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True))
        if not found:
            raise FileNotFoundError(f"no kernel_trace.csv under {path}")
        path = found[-1]
    rows = []
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            rows.append((int(r["Dispatch_Id"]), _short(r["Kernel_Name"]),
                         f"{r['Grid_Size_X']}x{r['Grid_Size_Y']}x{r['Grid_Size_Z']}",
                         int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    rows.sort()
    return rows, path


def _levels_of(rows):
    """Pyramid level of each dispatch of a run of SyN registrations: k_minmax
    (preprocessing) starts level 0 of a registration and k_resample_fields
    (the field adaptor) the next level; dispatches outside (calibration
    copies before the first registration) get None."""
    out, lv, started, prev_minmax = {}, None, False, False
    for did, k, grid, ns in rows:
        base = k.split("<")[0]
        if base == "k_minmax":
            if not prev_minmax:
                lv = 0
            started = True
        elif base == "k_resample_fields" and started:
            lv += 1
        prev_minmax = base == "k_minmax"
        out[did] = lv if started else None
    return out


def level_table(trace_dir: str, fetch_dir: str, write_dir: str, bench_json: str, copy_bytes: float) -> str:
    """Per kernel and pyramid level: rocprof mean duration, algorithmic bytes
    per launch (the engine's own count, from bench.py's per_level_kernels),
    achieved algorithmic GB/s, and HBM bytes per launch from the two counter
    passes (each scaled by the calibration copy kernel), against 8 TB/s."""
    import json
    line = None
    for ln in open(bench_json):
        ln = ln.strip()
        if ln.startswith("{") and '"per_level_kernels"' in ln:
            line = json.loads(ln)
    if line is None:
        raise RuntimeError(f"no bench.py JSON line with per_level_kernels in {bench_json}")
    plk = line["per_level_kernels"]
    level_names = list(plk.keys())
    cls_kernel = {"k_warp_pair": "k_warp_pair<3, float, false>", "k_cc3d": "k_cc3d<4, 16, 16>",
                  "k_smooth3d": "k_smooth3d<6, 16, 32, 3>", "k_compose": "k_compose<3>",
                  "k_inv_step": "k_inv_step<3, int>"}
    rows, tpath = _trace_rows(trace_dir)
    lvl = _levels_of(rows)
    dur = collections.defaultdict(list)
    for did, k, grid, ns in rows:
        if lvl.get(did) is not None:
            dur[(k, lvl[did])].append(ns)
    ctr = {}
    for cname, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        crow, cpath = counter_csv(d)
        trows, _ = _trace_rows(d)
        clvl = _levels_of(trows)
        kib = collections.defaultdict(float)
        n = collections.defaultdict(set)
        copy = []
        for did, k, grid, c, v in crow:
            if c != cname:
                continue
            if k.startswith("k_copy_calibration"):
                copy.append(v)
                continue
            if clvl.get(did) is not None:
                kib[(k, clvl[did])] += v
                n[(k, clvl[did])].add(did)
        scale = copy_bytes / (sum(copy) / len(copy) * 1024.0)
        ctr[cname] = {key: kib[key] / len(n[key]) * 1024.0 * scale for key in kib}
    out = [f"trace: {tpath}", "level grid      kernel                          launches  mean_us  alg_MB/launch  "
           "alg_GB/s  frac(8TB/s)  hbm_MB/launch  hbm_GB/s  hbm/alg"]
    for li, lname in enumerate(level_names):
        for cls, rec in plk[lname].items():
            kname = cls_kernel[cls]
            ds = dur.get((kname, li))
            if not ds:
                continue
            mean_us = sum(ds) / len(ds) / 1e3
            alg = rec["algorithmic_mb_per_launch"]
            gbs = alg * 1e6 / (mean_us * 1e-6) / 1e9
            hbm = (ctr["FETCH_SIZE"].get((kname, li), 0.0) + ctr["WRITE_SIZE"].get((kname, li), 0.0)) / 1e6
            hgbs = hbm * 1e6 / (mean_us * 1e-6) / 1e9
            out.append(f"{lname:14s}  {kname:30s}  {len(ds):8d}  {mean_us:7.1f}  {alg:13.2f}  {gbs:8.0f}  "
                       f"{gbs / 8000:11.3f}  {hbm:13.2f}  {hgbs:8.0f}  {hbm / alg if alg else 0:7.2f}")
    return "\n".join(out)


def trace_table(trace_dir: str) -> str:
    """Per kernel and launch shape of a kernel_trace.csv: calls, total ms,
    mean us and share of the kernel time."""
    rows, path = _trace_rows(trace_dir)
    agg = collections.OrderedDict()
    for did, k, grid, ns in rows:
        a = agg.setdefault((k, grid), [0, 0])
        a[0] += 1
        a[1] += ns
    total = sum(v[1] for v in agg.values())
    out = [f"trace: {path}", f"{'kernel':32s} {'grid(x,y,z)':18s} {'calls':>7s} {'total ms':>9s} {'mean us':>8s} {'%':>6s}"]
    for (k, grid), (n, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"{k:32s} {grid:18s} {n:7d} {ns / 1e6:9.3f} {ns / n / 1e3:8.2f} {100.0 * ns / total:6.1f}")
    out.append(f"total kernel time {total / 1e6:.3f} ms")
    return "\n".join(out)


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if argv and argv[0] == "trace":
        print(trace_table(argv[1]))
        return 0
    if argv and argv[0] == "levels":
        ap = argparse.ArgumentParser(prog="profiling levels")
        ap.add_argument("--trace", required=True)
        ap.add_argument("--fetch", required=True)
        ap.add_argument("--write", required=True)
        ap.add_argument("--bench-json", required=True, help="stdout of the traced bench.py run")
        ap.add_argument("--copy-bytes", type=float, required=True)
        a = ap.parse_args(argv[1:])
        print(level_table(a.trace, a.fetch, a.write, a.bench_json, a.copy_bytes))
        return 0
    if argv and argv[0] == "traffic":
        ap = argparse.ArgumentParser(prog="profiling traffic")
        ap.add_argument("--fetch", required=True)
        ap.add_argument("--write", required=True)
        ap.add_argument("--kernel", required=True)
        ap.add_argument("--copy-bytes", type=float, required=True)
        a = ap.parse_args(argv[1:])
        print(traffic(a.fetch, a.write, a.kernel, a.copy_bytes))
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("db", help="rocprofv3 results .db file, or a directory to search")
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    a = ap.parse_args(argv)
    db = a.db
    if os.path.isdir(db):
        found = sorted(glob.glob(os.path.join(db, "**", "*results.db"), recursive=True))
        if not found:
            print("no results.db under", db, file=sys.stderr)
            return 1
        db = found[-1]
    print(summarize(db, a.copy_bytes))
    return 0


if __name__ == "__main__":
    sys.exit(main())
