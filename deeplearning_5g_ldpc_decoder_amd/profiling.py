"""Reference workloads and summaries for rocprofv3 counters.

``python -m deeplearning_5g_ldpc_decoder_amd.profiling copy --mib 1024``
runs an elementwise device copy of a known size (one read and one write of
every byte) a few times.  Its FETCH_SIZE / WRITE_SIZE per dispatch, divided
into the known byte counts, gives the factor that converts the same counters
measured on the decode kernel into HBM bytes per launch.

``... profiling counters DIR --kernel NAME`` prints the mean per dispatch of
every counter in a ``rocprofv3 --pmc ... --output-format csv -d DIR`` run, for
the kernels whose name contains NAME.

``... profiling trace-summary DB`` prints the kernel table of a
``rocprofv3 --kernel-trace --stats`` database (calls, total and average
duration per kernel, and the duration spread of each) as markdown.

``... profiling valu-split DIR_A BENCH_A DIR_B BENCH_B`` splits SQ_INSTS_VALU
of the headline decode kernel, counted at two iteration limits (the bench
JSON lines of the same runs give each launch's executed iterations), into a
fixed part per codeblock and a part per executed codeblock-iteration.
"""
import argparse
import csv
import glob
import json
import os
import sys


def copy_workload(mib, reps):
    import torch
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)   # elementwise: reads n*4 bytes, writes n*4 bytes
    torch.cuda.synchronize()
    print(f"copy: {reps} dispatches of {n * 4} bytes read + {n * 4} bytes written",
          flush=True)


def counter_means(directory, kernel):
    """{counter: mean value per dispatch} over the dispatches of kernels whose
    name contains `kernel` (rocprofv3 CSV counter collection files)."""
    files = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {directory}")
    sums, dispatches = {}, {}
    for path in files:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if kernel not in row.get("Kernel_Name", ""):
                    continue
                name = row["Counter_Name"]
                key = (path, row.get("Dispatch_Id"))
                sums[name] = sums.get(name, 0.0) + float(row["Counter_Value"])
                dispatches.setdefault(name, set()).add(key)
    return {k: sums[k] / len(dispatches[k]) for k in sums}


def bench_iterations(path):
    """Total executed codeblock-iterations of one launch and its codeblock
    count, from a bench.py JSON line."""
    with open(path) as f:
        line = next(l for l in f if l.lstrip().startswith("{"))
    d = json.loads(line)
    n = d["config"]["global_batch"]
    return d["mean_iterations"] * n, n


def valu_split(dir_a, bench_a, dir_b, bench_b, kernel="ldpc_decode_fixed_kernel"):
    va = counter_means(dir_a, kernel)["SQ_INSTS_VALU"]
    vb = counter_means(dir_b, kernel)["SQ_INSTS_VALU"]
    ia, n = bench_iterations(bench_a)
    ib, nb = bench_iterations(bench_b)
    assert n == nb and abs(ia - ib) > 0.5 * n, "need two different iteration totals"
    per_it = (va - vb) / (ia - ib)
    per_cb = (va - per_it * ia) / n
    return {"kernel": kernel, "codeblocks_per_launch": n,
            "runs": [{"sq_insts_valu_per_dispatch": va, "codeblock_iterations": ia},
                     {"sq_insts_valu_per_dispatch": vb, "codeblock_iterations": ib}],
            "per_iteration_wave_instr": per_it, "per_codeblock_wave_instr": per_cb}


def trace_summary(db_path, top=8, by_grid=False):
    """Markdown table of the kernels of a rocprofv3 trace database (rocpd
    SQLite; durations in ns).  by_grid: one row per (kernel, workgroups per
    launch), which tells apart launches of one kernel with different grids
    (a persistent grid against one workgroup per tile)."""
    import sqlite3
    import statistics
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, duration, grid_x, workgroup_x from kernels").fetchall()
    per = {}
    for name, dur, gx, wx in rows:
        key = (name, gx // max(wx, 1) if by_grid else None)
        per.setdefault(key, []).append(dur / 1e3)
    total = sum(sum(v) for v in per.values())
    wg = " workgroups |" if by_grid else ""
    out = [f"| kernel |{wg} calls | total us | average us | median us | min us | max us | % |",
           "|---|" + ("---|" if by_grid else "") + "---|---|---|---|---|---|---|"]
    for (name, g), v in sorted(per.items(), key=lambda kv: -sum(kv[1]))[:top]:
        short = name if len(name) <= 110 else name[:107] + "..."
        gcol = f" {g} |" if by_grid else ""
        out.append(f"| `{short}` |{gcol} {len(v)} | {sum(v):.1f} | {statistics.mean(v):.2f} | "
                   f"{statistics.median(v):.2f} | {min(v):.2f} | {max(v):.2f} | "
                   f"{100 * sum(v) / total:.2f} |")
    return "\n".join(out)


def _kernel_rows(db_path, match):
    import sqlite3
    db = sqlite3.connect(db_path)
    return db.execute("select stream_id, start, end from kernels where name like ? "
                      "order by start", (f"%{match}%",)).fetchall()


def concurrency(rows):
    """Share of the time between the first start and the last end of rows
    (stream, start, end) during which k of them run at once, for each k."""
    events = sorted([(s, 1) for _, s, _ in rows] + [(e, -1) for _, _, e in rows])
    share, cur, last = {}, 0, events[0][0] if events else 0
    for t, d in events:
        share[cur] = share.get(cur, 0) + (t - last)
        last, cur = t, cur + d
    total = sum(share.values()) or 1
    return {k: v / total for k, v in sorted(share.items())}


def stream_gaps(db_path, match="ldpc_", min_kernels=20):
    """Markdown table, one row per stream with at least min_kernels kernels
    whose name contains match: kernels, median duration, median and 90th
    percentile idle gap between one such kernel's end and the next one's
    start on the stream, and the busy share of the stream's span; then the
    time share of 0, 1, 2, ... such kernels running at once over the trace."""
    import numpy as np
    rows = _kernel_rows(db_path, match)
    per = {}
    for sid, s, e in rows:
        per.setdefault(sid, []).append((s, e))
    out = ["| stream | kernels | median us | median gap us | p90 gap us | busy |",
           "|---|---|---|---|---|---|"]
    for sid, v in sorted(per.items()):
        if len(v) < min_kernels:
            continue
        dur = np.array([e - s for s, e in v]) / 1e3
        gap = np.array([v[i + 1][0] - v[i][1] for i in range(len(v) - 1)]) / 1e3
        span = (v[-1][1] - v[0][0]) / 1e3
        out.append(f"| {sid} | {len(v)} | {np.median(dur):.1f} | {np.median(gap):.1f} | "
                   f"{np.percentile(gap, 90):.1f} | {dur.sum() / span:.2f} |")
    conc = concurrency(rows)
    out.append("")
    out.append("Time share by kernels running at once: " +
               ", ".join(f"{k}: {v:.3f}" for k, v in conc.items()))
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--reps", type=int, default=3)
    k = sub.add_parser("counters")
    k.add_argument("dir")
    k.add_argument("--kernel", default="ldpc_decode")
    t = sub.add_parser("trace-summary")
    t.add_argument("db")
    t.add_argument("--by-grid", action="store_true",
                   help="one row per kernel and number of workgroups per launch")
    g = sub.add_parser("stream-gaps")
    g.add_argument("db")
    g.add_argument("--match", default="ldpc_",
                   help="kernels whose name contains this string")
    g.add_argument("--min-kernels", type=int, default=20)
    v = sub.add_parser("valu-split")
    v.add_argument("dir_a")
    v.add_argument("bench_a")
    v.add_argument("dir_b")
    v.add_argument("bench_b")
    args = ap.parse_args(argv)
    if args.cmd == "copy":
        copy_workload(args.mib, args.reps)
    elif args.cmd == "trace-summary":
        print(trace_summary(args.db, by_grid=args.by_grid))
    elif args.cmd == "stream-gaps":
        print(stream_gaps(args.db, args.match, args.min_kernels))
    elif args.cmd == "counters":
        print(json.dumps(counter_means(args.dir, args.kernel), indent=1))
    else:
        print(json.dumps(valu_split(args.dir_a, args.bench_a, args.dir_b, args.bench_b),
                         indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
