"""Profiling helpers for the benchmark: a plain device copy of known size
(the calibration kernel for the HBM byte counters), a kernel-trace summary
and the HBM-traffic figure from two rocprofv3 --pmc passes.

    python -m pypsi_amd.profiling copy [--mib 1024] [--reps 5]
    python -m pypsi_amd.profiling summary DB TITLE COMMAND
    python -m pypsi_amd.profiling traffic --copy-fetch DB --copy-write DB \
        --fetch DB --write DB --kernel NAME --workload W --out JSON

rocprofv3 writes SQLite result files (run_results.db); FETCH_SIZE and
WRITE_SIZE are per-dispatch kilobytes.  Each is scaled by the factor that
maps the copy kernel's reading onto its known bytes (read = write = size;
on MI355X FETCH_SIZE reads half of the copied bytes, WRITE_SIZE all).
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import sqlite3
import sys


def copy_kernel(mib: int = 1024, reps: int = 5) -> None:
    """A device-to-device copy of mib MiB, reps times (the HIP runtime's
    copyBuffer kernel: reads and writes each byte once)."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.arange(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy {mib} MiB x {reps}", flush=True)


def _db(path: str) -> str:
    if os.path.isdir(path):
        hits = sorted(glob.glob(os.path.join(path, "**", "*results.db"), recursive=True))
        if not hits:
            raise FileNotFoundError(f"no results.db under {path}")
        return hits[-1]
    return path


def summary(db: str, title: str, cmd: str) -> str:
    """Markdown table of a --kernel-trace run: per kernel calls, average and
    total duration, resources."""
    c = sqlite3.connect(_db(db))
    res = {r[0]: r[1:] for r in c.execute(
        "select distinct name, lds_size, scratch_size, vgpr_count, accum_vgpr_count, sgpr_count, grid_x, workgroup_x "
        "from kernels")}
    tot = c.execute("select sum(duration) from kernels").fetchone()[0] or 1
    out = [f"# {title}", "", "Command (from the repository root on the GPU host):", "", "```", cmd, "```", "",
           "| kernel | calls | avg us | min us | max us | total us | % | VGPR | AGPR | SGPR | LDS B | scratch B/lane "
           "| grid | block |", "|---|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for name, n, avg, mn, mx, s in c.execute(
            "select name, count(*), avg(duration), min(duration), max(duration), sum(duration) from kernels "
            "group by name order by sum(duration) desc"):
        lds, scr, vg, ag, sg, gx, wx = res[name]
        short = name.split("(")[0].replace("void ", "")[:64]
        out.append(f"| {short} | {n} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | {s / 1e3:.1f} | "
                   f"{100 * s / tot:.2f} | {vg} | {ag} | {sg} | {lds} | {scr} | {gx} | {wx} |")
    return "\n".join(out)


def counter_per_dispatch(db: str, counter: str, kernel: str) -> list[float]:
    """Per-dispatch values of a counter for kernels whose name contains
    `kernel`."""
    c = sqlite3.connect(_db(db))
    rows = c.execute("select dispatch_id, kernel_name, sum(value) from counters_collection where counter_name = ? "
                     "group by dispatch_id order by dispatch_id", (counter,)).fetchall()
    return [v for _, name, v in rows if kernel in name]


def traffic(copy_fetch: str, copy_write: str, fetch: str, write: str, kernel: str, copy_mib: int) -> dict:
    size_kib = copy_mib * 1024.0
    cf = sorted(counter_per_dispatch(copy_fetch, "FETCH_SIZE", "copyBuffer"))
    cw = sorted(counter_per_dispatch(copy_write, "WRITE_SIZE", "copyBuffer"))
    f_fetch = size_kib / cf[len(cf) // 2]
    f_write = size_kib / cw[len(cw) // 2]
    kf = counter_per_dispatch(fetch, "FETCH_SIZE", kernel)
    kw = counter_per_dispatch(write, "WRITE_SIZE", kernel)
    fb = f_fetch * 1024.0 * sum(kf) / len(kf)
    wb = f_write * 1024.0 * sum(kw) / len(kw)
    return {"kernel": kernel, "fetch_bytes": fb, "write_bytes": wb, "bytes": fb + wb, "dispatches": len(kf),
            "copy_fetch_factor": f_fetch, "copy_write_factor": f_write, "copy_mib": copy_mib}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m pypsi_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("copy")
    p.add_argument("--mib", type=int, default=1024)
    p.add_argument("--reps", type=int, default=5)
    p = sub.add_parser("summary")
    p.add_argument("db")
    p.add_argument("title")
    p.add_argument("command")
    p = sub.add_parser("traffic")
    for k in ("copy-fetch", "copy-write", "fetch", "write", "kernel", "workload", "out"):
        p.add_argument("--" + k, required=True)
    p.add_argument("--copy-mib", type=int, default=1024)
    p.add_argument("--label", default="fast")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_kernel(a.mib, a.reps)
    elif a.cmd == "summary":
        print(summary(a.db, a.title, a.command))
    else:
        t = traffic(a.copy_fetch, a.copy_write, a.fetch, a.write, a.kernel, a.copy_mib)
        data = {}
        if os.path.exists(a.out):
            with open(a.out) as f:
                data = json.load(f)
        data.setdefault(a.workload, {})[a.label] = t
        with open(a.out, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        print(json.dumps(t))
    return 0


if __name__ == "__main__":
    sys.exit(main())
