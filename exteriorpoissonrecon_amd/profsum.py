"""Summaries of rocprofv3 CSV output for the profiles/ folder.

    python -m exteriorpoissonrecon_amd.profsum stats <dir> [--top N] [--kernel NAME]
        per-kernel table from a `rocprofv3 --kernel-trace --stats --output-format csv` run;
        with --kernel, also the duration of the largest-grid dispatches of NAME
        (the finest level) from the per-dispatch trace
    python -m exteriorpoissonrecon_amd.profsum traffic <fetch_dir> <write_dir> --kernel NAME --copy-bytes B
        HBM bytes per dispatch of the largest-grid dispatches of kernel NAME from
        two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs, scaled by what
        the same counters read for the KCopy calibration kernel (B bytes read and
        B bytes written per dispatch, launched by bench.py through epr_time_copy)
    python -m exteriorpoissonrecon_amd.profsum counters <dir> --kernel NAME [--nodes N]
        mean per-dispatch value of every counter of a `rocprofv3 --pmc` run over the
        largest-grid dispatches of NAME; with --nodes, also per node (N nodes a launch):
        an SQ_INSTS_* counter counts wavefront instructions, so its per-node figure is
        scaled by the 64 lanes of a wavefront (lane operations per node; SQ_WAVES
        becomes lanes per node)

Counter values are KiB per dispatch.  The scale factor of a counter is
(known KiB of the copy) / (counter value of the copy); the summary prints it.
"""
import argparse
import csv
import glob
import os
import sys
from collections import defaultdict


def _files(d, suffix):
    return sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))


def _short(name, width=90):
    name = name.replace("void ", "")
    return name if len(name) <= width else name[:width - 3] + "..."


def stats(d, top=25):
    rows = []
    for f in _files(d, "kernel_stats.csv"):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    if not rows:
        raise SystemExit(f"no kernel_stats.csv under {d}")
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{100 * float(r['TotalDurationNs']) / total:.1f} |")
    out.append(f"\nTotal kernel time {total / 1e6:.3f} ms over {sum(int(r['Calls']) for r in rows)} dispatches.")
    return "\n".join(out)


def finest(d, kernel):
    """Durations of the largest-grid dispatches of `kernel` in the kernel trace."""
    rows = []
    for f in _files(d, "kernel_trace.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if kernel in r["Kernel_Name"]:
                    g = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
                    rows.append((g, int(r["End_Timestamp"]) - int(r["Start_Timestamp"]), r))
    if not rows:
        raise SystemExit(f"no dispatches of {kernel} under {d}")
    gmax = max(g for g, _, _ in rows)
    big = [(g, t, r) for g, t, r in rows if g * 4 >= gmax]
    out = []
    # one line per workgroup size (an autotuner's trial launches use other sizes than the picked plan)
    for wg in sorted({r["Workgroup_Size_X"] for _, _, r in big}, key=int):
        sel = [(g, t, r) for g, t, r in big if r["Workgroup_Size_X"] == wg]
        ts = sorted(t for _, t, _ in sel)
        r0 = sel[0][2]
        out.append(f"{kernel}, largest-grid dispatches (grid sizes {sorted({g for g, _, _ in sel})}), workgroup {wg}: "
                   f"{len(ts)} dispatches, avg {sum(ts) / len(ts) / 1e3:.1f} us, median {ts[len(ts) // 2] / 1e3:.1f} us, "
                   f"min {ts[0] / 1e3:.1f} us, max {ts[-1] / 1e3:.1f} us; "
                   f"VGPR {r0['VGPR_Count']}, AGPR {r0['Accum_VGPR_Count']}, SGPR {r0['SGPR_Count']}, "
                   f"LDS {r0['LDS_Block_Size']} B, scratch {r0['Scratch_Size']} B")
    return "\n\n".join(out)


def _counters(d, counter):
    """(kernel name, grid size) -> list of counter values (one per dispatch)."""
    vals = defaultdict(list)
    for f in _files(d, "counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                vals[(r["Kernel_Name"], int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    if not vals:
        raise SystemExit(f"no {counter} rows under {d}")
    return vals


def _select(vals, kernel):
    keys = [k for k in vals if kernel in k[0]]
    if not keys:
        raise SystemExit(f"no dispatches of {kernel}")
    gmax = max(k[1] for k in keys)
    big = [k for k in keys if k[1] * 4 >= gmax]          # the largest-grid dispatches (the finest level)
    v = [x for k in big for x in vals[k]]
    return sum(v) / len(v), len(v), sorted({k[1] for k in big})


def traffic(fetch_dir, write_dir, kernel, copy_bytes):
    fv, wv = _counters(fetch_dir, "FETCH_SIZE"), _counters(write_dir, "WRITE_SIZE")
    cf, _, _ = _select(fv, "KCopy")
    cw, _, _ = _select(wv, "KCopy")
    known_kib = copy_bytes / 1024.0
    sf, sw = known_kib / cf, known_kib / cw
    kf, nf, grids = _select(fv, kernel)
    kw, nw, _ = _select(wv, kernel)
    bytes_per = (kf * sf + kw * sw) * 1024.0
    out = [
        f"calibration KCopy ({copy_bytes} B read + {copy_bytes} B written per dispatch): "
        f"FETCH_SIZE {cf:.1f} KiB -> factor {sf:.4f}; WRITE_SIZE {cw:.1f} KiB -> factor {sw:.4f}",
        f"{kernel}: {nf} dispatches (grid sizes {grids}), mean FETCH_SIZE {kf:.1f} KiB, "
        f"mean WRITE_SIZE {kw:.1f} KiB",
        f"traffic_bytes_per_launch: {bytes_per:.6g}",
        f"fetch_bytes_per_launch: {kf * sf * 1024.0:.6g}",
        f"write_bytes_per_launch: {kw * sw * 1024.0:.6g}",
    ]
    return "\n".join(out)


def counters(d, kernel, nodes=0):
    names = set()
    for f in _files(d, "counter_collection.csv"):
        with open(f) as fh:
            names |= {r["Counter_Name"] for r in csv.DictReader(fh)}
    if not names:
        raise SystemExit(f"no counter rows under {d}")
    out = [f"{kernel}: mean per dispatch over the largest-grid dispatches" + (f" ({nodes} nodes each)" if nodes else ""),
           "", "| counter | dispatches | grid sizes | per dispatch |" + (" x 64 lanes / node |" if nodes else ""),
           "|---|---|---|---|" + ("---|" if nodes else "")]
    for c in sorted(names):
        v, n, g = _select(_counters(d, c), kernel)
        out.append(f"| {c} | {n} | {g} | {v:.4g} |" + (f" {v * 64 / nodes:.4g} |" if nodes else ""))
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=25)
    s.add_argument("--kernel", action="append", default=[])
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--kernel", required=True)
    c.add_argument("--nodes", type=int, default=0)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats(a.dir, a.top))
        for k in a.kernel:
            print("\n" + finest(a.dir, k))
    elif a.cmd == "counters":
        print(counters(a.dir, a.kernel, a.nodes))
    else:
        print(traffic(a.fetch_dir, a.write_dir, a.kernel, a.copy_bytes))


if __name__ == "__main__":
    sys.exit(main())
