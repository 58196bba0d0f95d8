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


def traffic(fetch_dir: str, write_dir: str, kernel: str, copy_bytes: float) -> str:
    lines = []
    per = {}
    for cname, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows, path = counter_csv(d)
        lines.append(f"{cname}: {path}")
        kib = collections.defaultdict(float)  # (kernel, grid) -> summed KiB over dispatches
        disp = collections.defaultdict(set)
        for did, k, grid, ctr, v in rows:
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


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
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
