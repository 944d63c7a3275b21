"""Summaries of rocprofv3 CSV output for the worm step kernels.

  python -m nwamd.profsum trace <kernel_trace.csv> [--last N] [--by-grid [--fold-below G]]
      per-kernel dispatch durations (markdown table); --by-grid keeps the
      launches of one kernel over different grid sizes (work-items) apart,
      --fold-below G merges each kernel's launches over grids below G
  python -m nwamd.profsum traffic <FETCH_SIZE counter csv> <WRITE_SIZE counter csv> --copy NAME --copy-bytes B
      HBM bytes per dispatch from the two counter passes, each scaled by what
      the same counter reads for a copy kernel of known size (NAME: a
      substring of its kernel name; it reads B bytes and writes B bytes);
      prints markdown and, with --json PATH, writes {kernel: bytes} there

  python -m nwamd.profsum issue <SQ counter csv> [--cus 256] [--ghz 2.4] [--min-calls 5]
      instructions per wavefront (SQ_INSTS_VALU / SALU / LDS / VMEM / SMEM
      over SQ_WAVES) and the time the dispatch's VALU and SALU instructions
      take to issue at one per cycle per CU (a CU issues at most one VALU
      wave-instruction per cycle over its four SIMDs, and its one scalar unit
      one SALU instruction per cycle): the instruction-issue floor of a kernel

The traffic counters are per dispatch in KiB.  FETCH_SIZE counts L2 fills
from the fabric (HBM and the MALL), WRITE_SIZE L2 write-backs.
"""
import argparse
import csv
import json
import sys
from collections import defaultdict


def kernel_key(name):
    """Short kernel name: template arguments kept, namespace and signature dropped."""
    name = name.replace("void ", "")
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return name[:cut].replace("nw::", "")


def trace(path, last=0, by_grid=False, fold_below=0):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    if last:
        rows = rows[-last:]
    agg = defaultdict(list)
    for r in rows:
        key = kernel_key(r["Kernel_Name"])
        if by_grid:
            g = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
            key = (key, f"< {fold_below}" if g < fold_below else str(g))
        agg[key].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    tot = sum(sum(v) for v in agg.values())
    head = "| kernel | grid | calls |" if by_grid else "| kernel | calls |"
    out = [head + " avg us | min us | max us | total ms | % |",
           "|---|" + ("---:|" if by_grid else "") + "---:|---:|---:|---:|---:|---:|"]
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        name = f"`{k[0]}` | {k[1]}" if by_grid else f"`{k}`"
        out.append(f"| {name} | {len(v)} | {sum(v) / len(v) / 1e3:.2f} | {min(v) / 1e3:.2f} | {max(v) / 1e3:.2f} | "
                   f"{sum(v) / 1e6:.3f} | {100 * sum(v) / tot:.1f} |")
    out.append(f"\nTotal kernel time {tot / 1e6:.3f} ms over {len(rows)} dispatches.")
    return "\n".join(out)


def counter_means(path, counter):
    agg = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            agg[kernel_key(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def traffic(fetch_csv, write_csv, copy, copy_bytes):
    fetch, write = counter_means(fetch_csv, "FETCH_SIZE"), counter_means(write_csv, "WRITE_SIZE")
    ck = [k for k in fetch if copy in k]
    if not ck:
        raise SystemExit(f"no kernel matching {copy!r} in {fetch_csv}")
    ck = ck[0]
    f_scale = copy_bytes / (fetch[ck] * 1024.0)
    w_scale = copy_bytes / (write[ck] * 1024.0)
    res = {}
    lines = [f"Calibration: `{ck}` reads and writes {copy_bytes} B; FETCH_SIZE reads "
             f"{fetch[ck]:.0f} KiB (scale {f_scale:.3f}), WRITE_SIZE {write[ck]:.0f} KiB (scale {w_scale:.3f}).\n",
             "| kernel | fetch MB | write MB | traffic MB per dispatch |", "|---|---:|---:|---:|"]
    for k in sorted(fetch, key=lambda k: -(fetch[k] + write.get(k, 0.0))):
        if k == ck or k not in write:
            continue
        fb, wb = fetch[k] * 1024.0 * f_scale, write[k] * 1024.0 * w_scale
        res[k] = fb + wb
        lines.append(f"| `{k}` | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {(fb + wb) / 1e6:.1f} |")
    return "\n".join(lines), res, dict(fetch_scale=f_scale, write_scale=w_scale, copy_kernel=ck)


def issue(path, cus, ghz, min_calls):
    per = defaultdict(lambda: defaultdict(list))
    for r in csv.DictReader(open(path)):
        per[kernel_key(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    lines = ["| kernel | calls | waves | VALU / wave | SALU / wave | LDS / wave | VMEM / wave | SMEM / wave | "
             "VALU issue us | SALU issue us |", "|---|" + "---:|" * 9]
    rows = []
    for k, c in per.items():
        n = len(c.get("SQ_WAVES", []))
        if n < min_calls or "SQ_INSTS_VALU" not in c:
            continue
        m = {name: sum(v) / len(v) for name, v in c.items()}
        w = m["SQ_WAVES"]
        if w <= 0:
            continue
        rate = cus * ghz * 1e3  # instructions per microsecond over the chip
        rows.append((m["SQ_INSTS_VALU"] / rate, k, n, w, m))
    for t, k, n, w, m in sorted(rows, reverse=True):
        lines.append(f"| `{k}` | {n} | {w:.0f} | {m['SQ_INSTS_VALU'] / w:.0f} | {m['SQ_INSTS_SALU'] / w:.0f} | "
                     f"{m.get('SQ_INSTS_LDS', 0) / w:.1f} | {m.get('SQ_INSTS_VMEM', 0) / w:.1f} | "
                     f"{m.get('SQ_INSTS_SMEM', 0) / w:.1f} | {t:.2f} | "
                     f"{m['SQ_INSTS_SALU'] / (cus * ghz * 1e3):.2f} |")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m nwamd.profsum")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--last", type=int, default=0, help="only the last N dispatches")
    t.add_argument("--by-grid", action="store_true", help="one row per kernel and grid size")
    t.add_argument("--fold-below", type=int, default=0, help="with --by-grid: one row for all grids below this")
    p = sub.add_parser("traffic")
    p.add_argument("fetch_csv")
    p.add_argument("write_csv")
    p.add_argument("--copy", required=True)
    p.add_argument("--copy-bytes", type=float, required=True)
    p.add_argument("--json")
    q = sub.add_parser("issue")
    q.add_argument("csv")
    q.add_argument("--cus", type=int, default=256)
    q.add_argument("--ghz", type=float, default=2.4)
    q.add_argument("--min-calls", type=int, default=5)
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print(trace(a.csv, a.last, a.by_grid, a.fold_below))
    elif a.cmd == "issue":
        print(issue(a.csv, a.cus, a.ghz, a.min_calls))
    else:
        text, res, cal = traffic(a.fetch_csv, a.write_csv, a.copy, a.copy_bytes)
        print(text)
        if a.json:
            json.dump(dict(bytes_per_dispatch=res, calibration=cal), open(a.json, "w"), indent=1)


if __name__ == "__main__":
    sys.exit(main())
