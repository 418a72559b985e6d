"""Turn rocprofv3 CSV output into the short summaries kept under profiles/.

  python -m helios_amd.profiling stats  <kernel_stats.csv> --cmd "<command>" --out profiles/<name>.md
  python -m helios_amd.profiling traffic <fetch counter_collection.csv> <write counter_collection.csv>
         --kernel k_extend --cmd "<command>" --out profiles/extend_traffic.json

`traffic` scales FETCH_SIZE and WRITE_SIZE (KiB per dispatch) by what the same
counters read for the plain copy kernel k_copy_probe, which moves a known
1 GiB in and 1 GiB out per dispatch.
"""
import argparse
import collections
import csv
import json
from pathlib import Path

PROBE_BYTES = 1 << 30


def short(name):
    n = name.split("(")[0].replace("void ", "")
    return n if len(n) < 90 else n[:87] + "..."


def stats_md(csv_path, cmd, top=12):
    rows = list(csv.DictReader(open(csv_path)))
    lines = [f"Command: `{cmd}`", "", "| kernel | calls | total ms | avg us | min us | max us | % |",
             "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        lines.append(f"| {short(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                     f"{float(r['Percentage']):.2f} |")
    return "\n".join(lines) + "\n"


def per_dispatch(csv_path):
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(csv_path)):
        agg[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return agg


def traffic(fetch_csv, write_csv, kernel, cmd):
    f, w = per_dispatch(fetch_csv), per_dispatch(write_csv)
    probe_f = [v for k, vs in f.items() if "k_copy_probe" in k for v in vs]
    probe_w = [v for k, vs in w.items() if "k_copy_probe" in k for v in vs]
    f_scale = PROBE_BYTES / (sum(probe_f) / len(probe_f) * 1024)
    w_scale = PROBE_BYTES / (sum(probe_w) / len(probe_w) * 1024)
    kf = [v for k, vs in f.items() if kernel in k and "true" not in k for v in vs]
    kw = [v for k, vs in w.items() if kernel in k and "true" not in k for v in vs]
    fetch_b = sum(kf) / len(kf) * 1024 * f_scale
    write_b = sum(kw) / len(kw) * 1024 * w_scale
    return {
        "kernel": kernel, "dispatches": len(kf), "command": cmd,
        "fetch_kib_per_dispatch_raw": round(sum(kf) / len(kf), 1),
        "write_kib_per_dispatch_raw": round(sum(kw) / len(kw), 1),
        "copy_probe_fetch_kib_raw": round(sum(probe_f) / len(probe_f), 1),
        "copy_probe_write_kib_raw": round(sum(probe_w) / len(probe_w), 1),
        "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
        "hbm_read_bytes_per_launch": round(fetch_b), "hbm_write_bytes_per_launch": round(write_b),
        "hbm_bytes_per_launch": round(fetch_b + write_b),
    }


def main():
    p = argparse.ArgumentParser()
    sub = p.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--cmd", required=True)
    s.add_argument("--title", default="")
    s.add_argument("--out", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--kernel", default="k_extend")
    t.add_argument("--cmd", required=True)
    t.add_argument("--out", required=True)
    a = p.parse_args()
    if a.what == "stats":
        text = (f"# {a.title}\n\n" if a.title else "") + stats_md(a.csv, a.cmd)
        Path(a.out).write_text(text)
    else:
        Path(a.out).write_text(json.dumps(traffic(a.fetch_csv, a.write_csv, a.kernel, a.cmd), indent=1) + "\n")
    print("wrote", a.out)


if __name__ == "__main__":
    main()
