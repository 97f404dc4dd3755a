"""Summaries of rocprofv3 CSV output for the render kernels.

    python -m axiomr_amd.rocprof_summary stats DIR [--title T] [--command C]
        Markdown table of kernel_stats.csv (calls, average/min/max duration).

    python -m axiomr_amd.rocprof_summary sq DIR [--stats DIR]
        Per-kernel lane utilisation, VALU instructions per wave and VALU busy
        from one --pmc pass of SQ_COUNTERS (and the durations of a stats run).

    python -m axiomr_amd.rocprof_summary traffic --fetch DIR --write DIR \
        --copy-fetch DIR --copy-write DIR --copy-bytes N [--tunables JSON]
        [--workload W] [--command C ...] [--out JSON]
        Per-kernel HBM bytes per dispatch from two --pmc passes (FETCH_SIZE,
        then WRITE_SIZE, both in KiB per dispatch), each scaled by the factor
        that makes the same counter read N bytes for a plain device copy of
        N bytes.

Kernel names are shortened to the function name plus template arguments
(`k_raster<PhongPoint,256,false>`).
"""
from __future__ import annotations

import argparse
import csv
import json
import re
import statistics
import sys
from collections import defaultdict
from pathlib import Path
from typing import Dict, List


def short_name(name: str) -> str:
    m = re.search(r"(k_\w+)(<[^()]*>)?", name)
    if m:
        tmpl = (m.group(2) or "").replace("axr::", "").replace(" ", "")
        return m.group(1) + tmpl
    name = re.sub(r"\(.*", "", name)
    return name[-80:]


def _find(dirpath: Path, suffix: str) -> List[Path]:
    files = sorted(Path(dirpath).rglob(f"*{suffix}"))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {dirpath}")
    return files


def _col(row: Dict[str, str], *names: str) -> str:
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(f"none of {names} in {list(row)}")


def kernel_stats(dirpath: Path) -> List[Dict]:
    rows = []
    for f in _find(dirpath, "kernel_stats.csv"):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                rows.append({
                    "name": short_name(_col(r, "Name", "KernelName")),
                    "calls": int(_col(r, "Calls")),
                    "total_ns": float(_col(r, "TotalDurationNs")),
                    "avg_ns": float(_col(r, "AverageNs")),
                    "min_ns": float(_col(r, "MinNs")),
                    "max_ns": float(_col(r, "MaxNs")),
                    "pct": float(_col(r, "Percentage")),
                })
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def stats_markdown(dirpath: Path, title: str = "", command: str = "") -> str:
    out = []
    if title:
        out.append(f"# {title}\n")
    if command:
        out.append("Command:\n\n```\n" + command + "\n```\n")
    out.append("| kernel | calls | avg us | min us | max us | % of GPU time |")
    out.append("|---|---:|---:|---:|---:|---:|")
    for r in kernel_stats(dirpath):
        out.append(f"| `{r['name']}` | {r['calls']} | {r['avg_ns'] / 1e3:.2f} | {r['min_ns'] / 1e3:.2f} | "
                   f"{r['max_ns'] / 1e3:.2f} | {r['pct']:.2f} |")
    return "\n".join(out) + "\n"


def counter_per_kernel(dirpath: Path, counter: str) -> Dict[str, Dict]:
    """Average (and count) of one counter per dispatch, by kernel."""
    vals: Dict[str, List[float]] = defaultdict(list)
    for f in _find(dirpath, "counter_collection.csv"):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if _col(r, "Counter_Name").strip() != counter:
                    continue
                vals[short_name(_col(r, "Kernel_Name"))].append(float(_col(r, "Counter_Value")))
    return {k: {"avg": statistics.fmean(v), "median": statistics.median(v), "dispatches": len(v)}
            for k, v in vals.items()}


def copy_factor(dirpath: Path, counter: str, nbytes: int) -> Dict:
    """Scale factor bytes/counter-unit from the largest kernel of a plain
    copy run (the copy kernel reads and writes `nbytes` each)."""
    per = counter_per_kernel(dirpath, counter)
    name, rec = max(per.items(), key=lambda kv: kv[1]["median"])
    return {"kernel": name, "reading": rec["median"], "bytes": nbytes, "factor": nbytes / rec["median"]}


def traffic(fetch_dir: Path, write_dir: Path, copy_fetch: Path, copy_write: Path, copy_bytes: int) -> Dict:
    cf = copy_factor(copy_fetch, "FETCH_SIZE", copy_bytes)
    cw = copy_factor(copy_write, "WRITE_SIZE", copy_bytes)
    fetch = counter_per_kernel(fetch_dir, "FETCH_SIZE")
    write = counter_per_kernel(write_dir, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        fk = fetch.get(k, {}).get("median", 0.0)
        wk = write.get(k, {}).get("median", 0.0)
        kernels[k] = {"fetch_kib": fk, "write_kib": wk, "fetch_bytes": fk * cf["factor"],
                      "write_bytes": wk * cw["factor"], "traffic_bytes": fk * cf["factor"] + wk * cw["factor"],
                      "dispatches": max(fetch.get(k, {}).get("dispatches", 0), write.get(k, {}).get("dispatches", 0))}
    return {"copy_fetch": cf, "copy_write": cw, "kernels": kernels}


SQ_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES",
               "SQ_BUSY_CYCLES", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT")


def sq_table(sq_dir: Path, stats_dir: Path, clock_ghz: float = 2.4, simds: int = 1024) -> List[Dict]:
    """Per kernel, from one --pmc pass of SQ_COUNTERS (medians over the
    kernel's dispatches, whole device) and the kernel's average duration
    from a --kernel-trace --stats run of the same command:
      lane utilisation = SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU)
          (the share of the 64 lanes that did work in the VALU cycles),
      VALU busy = 4 x SQ_ACTIVE_INST_VALU / (SIMDs x kernel cycles),
      VALU instructions per wave = SQ_INSTS_VALU / SQ_WAVES."""
    per = {c: counter_per_kernel(sq_dir, c) for c in SQ_COUNTERS}
    dur = {r["name"]: r["avg_ns"] for r in kernel_stats(stats_dir)} if stats_dir else {}
    rows = []
    for k in sorted(per["SQ_WAVES"], key=lambda n: -per["SQ_INSTS_VALU"].get(n, {}).get("median", 0.0)):
        m = {c: per[c].get(k, {}).get("median", 0.0) for c in SQ_COUNTERS}
        act = m["SQ_ACTIVE_INST_VALU"]
        ns = dur.get(k)
        rows.append({
            "kernel": k, "waves": m["SQ_WAVES"],
            "valu_per_wave": m["SQ_INSTS_VALU"] / m["SQ_WAVES"] if m["SQ_WAVES"] else 0.0,
            "lane_util": m["SQ_THREAD_CYCLES_VALU"] / (64.0 * act) if act else 0.0,
            "valu_busy": 4.0 * act / (simds * ns * clock_ghz) if ns else None,
            "avg_us": ns / 1e3 if ns else None,
            "lds_insts": m["SQ_INSTS_LDS"], "lds_bank_conflicts": m["SQ_LDS_BANK_CONFLICT"],
            "counters": m,
        })
    return rows


def sq_markdown(rows: List[Dict]) -> str:
    out = ["| kernel | avg us | waves | VALU inst / wave | lane utilisation | VALU busy | LDS inst | LDS bank-conflict cycles |",
           "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        if not r["counters"]["SQ_INSTS_VALU"]:
            continue
        us = f"{r['avg_us']:.2f}" if r["avg_us"] is not None else "-"
        busy = f"{100 * r['valu_busy']:.1f} %" if r["valu_busy"] is not None else "-"
        out.append(f"| `{r['kernel']}` | {us} | {r['waves']:.0f} | {r['valu_per_wave']:.0f} | "
                   f"{r['lane_util']:.2f} | {busy} | {r['lds_insts']:.3g} | {r['lds_bank_conflicts']:.3g} |")
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--title", default="")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    for a in ("--fetch", "--write", "--copy-fetch", "--copy-write"):
        t.add_argument(a, required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--tunables", default="", help="JSON of the axr_config the profiled run used")
    t.add_argument("--workload", default="")
    t.add_argument("--command", action="append", default=[], help="profiling commands, recorded verbatim")
    t.add_argument("--out", default="")
    q = sub.add_parser("sq")
    q.add_argument("dir", help="output of one --pmc pass of SQ_COUNTERS")
    q.add_argument("--stats", default="", help="output of a --kernel-trace --stats run of the same command")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        sys.stdout.write(stats_markdown(Path(args.dir), args.title, args.command))
    elif args.cmd == "sq":
        sys.stdout.write(sq_markdown(sq_table(Path(args.dir), Path(args.stats) if args.stats else None)))
    else:
        res = traffic(Path(args.fetch), Path(args.write), Path(args.copy_fetch), Path(args.copy_write),
                      args.copy_bytes)
        if args.tunables:
            res["tunables"] = json.loads(args.tunables)
        if args.workload:
            res["workload"] = args.workload
        if args.command:
            res["commands"] = args.command
        txt = json.dumps(res, indent=1, sort_keys=True)
        if args.out:
            Path(args.out).write_text(txt + "\n")
        print(txt)


if __name__ == "__main__":
    main()
