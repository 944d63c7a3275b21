"""Summaries of rocprofv3 CSV output for the worm step kernels.

  python -m nwamd.profsum trace <kernel_trace.csv> [--last N]
      per-kernel dispatch durations (markdown table)
  python -m nwamd.profsum traffic <FETCH_SIZE counter csv> <WRITE_SIZE counter csv> --copy NAME --copy-bytes B
      HBM bytes per dispatch from the two counter passes, each scaled by what
      the same counter reads for a copy kernel of known size (NAME: a
      substring of its kernel name; it reads B bytes and writes B bytes);
      prints markdown and, with --json PATH, writes {kernel: bytes} there

The counters are per dispatch in KiB.  FETCH_SIZE counts L2 fills from the
fabric (HBM and the MALL), WRITE_SIZE L2 write-backs.
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


def trace(path, last=0):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    if last:
        rows = rows[-last:]
    agg = defaultdict(list)
    for r in rows:
        agg[kernel_key(r["Kernel_Name"])].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    tot = sum(sum(v) for v in agg.values())
    out = ["| kernel | calls | avg us | min us | max us | total ms | % |", "|---|---:|---:|---:|---:|---:|---:|"]
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| `{k}` | {len(v)} | {sum(v) / len(v) / 1e3:.2f} | {min(v) / 1e3:.2f} | {max(v) / 1e3:.2f} | "
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


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m nwamd.profsum")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--last", type=int, default=0, help="only the last N dispatches")
    p = sub.add_parser("traffic")
    p.add_argument("fetch_csv")
    p.add_argument("write_csv")
    p.add_argument("--copy", required=True)
    p.add_argument("--copy-bytes", type=float, required=True)
    p.add_argument("--json")
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print(trace(a.csv, a.last))
    else:
        text, res, cal = traffic(a.fetch_csv, a.write_csv, a.copy, a.copy_bytes)
        print(text)
        if a.json:
            json.dump(dict(bytes_per_dispatch=res, calibration=cal), open(a.json, "w"), indent=1)


if __name__ == "__main__":
    sys.exit(main())
