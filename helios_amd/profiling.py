"""Turn rocprofv3 CSV output into the short summaries kept under profiles/.

  python -m helios_amd.profiling stats  <kernel_stats.csv> --cmd "<command>" --out profiles/<name>.md
         [--trace <kernel_trace.csv> [--timeline <pass>]] [--title ...] [--note ...]
  python -m helios_amd.profiling traffic <fetch counter_collection.csv> <write counter_collection.csv>
         --kernel k_extend --cmd "<command>" --out profiles/extend_traffic.json
  python -m helios_amd.profiling mfma <kernel_stats.csv> <counter_collection.csv> --cmd "<commands>"
         --out profiles/<name>.md
  python -m helios_amd.profiling issue <counter_collection.csv> --cmd "<command>" --shape '<json>'
         --out profiles/extend_issue.json

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


def demangle(name):
    """Readable form of the denoiser's mangled template kernel names
    (_ZN2hh2dn9k_conv3x3ILi48ELi2ELi32ELi32ELi0EE... -> hh::dn::k_conv3x3<48, 2, 32, 32, 0>)."""
    import re
    m = re.match(r"_ZN2hh2dn(\d+)", name)
    if not m:
        return name
    n = int(m.group(1))
    rest = name[m.end():]
    kname, tail = rest[:n], rest[n:]
    args = re.match(r"I((?:Li-?\d+E)+)E", tail)
    if not args:
        return "hh::dn::" + kname
    return f"hh::dn::{kname}<{', '.join(re.findall(r'Li(-?[0-9]+)E', args.group(1)))}>"


def short(name):
    name = demangle(name)
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


def passes_md(trace_csv, max_rows=12):
    """Extension launches (k_extend_packet + k_extend, not the counting
    instantiations) grouped by wavefront pass: a pass starts at each
    k_generate dispatch. Lists launches and average duration per pass, in
    dispatch order, so the timed pass can be matched with the bench's
    HIP-event figure."""
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    passes, cur = [], None
    for r in rows:
        n = r["Kernel_Name"]
        if "k_generate<" in n:
            cur = {"gen": short(n), "ext": []}
            passes.append(cur)
        elif cur is not None and ("k_extend_packet<" in n or ("k_extend<" in n and "false>" in n)):
            cur["ext"].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    lines = ["", "Extension launches per wavefront pass (a pass starts at each k_generate dispatch):", "",
             "| pass | generate | extension launches | avg us |", "|---|---|---|---|"]
    for i, pz in enumerate(passes[:max_rows]):
        if pz["ext"]:
            lines.append(f"| {i + 1} | {pz['gen']} | {len(pz['ext'])} | {sum(pz['ext']) / len(pz['ext']) / 1e3:.1f} |")
    return "\n".join(lines) + "\n"


def timeline_md(trace_csv, pass_no):
    """Every dispatch of wavefront pass `pass_no` (1-based, a pass starts at
    each k_generate dispatch) with its hardware queue, start and end relative
    to the pass's first dispatch, and duration: shows which launches ran side
    by side (the shadow launch beside the next extension launch)."""
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    starts = [i for i, r in enumerate(rows) if "k_generate<" in r["Kernel_Name"]]
    if pass_no < 1 or pass_no > len(starts):
        return ""
    i0 = starts[pass_no - 1]
    i1 = starts[pass_no] if pass_no < len(starts) else len(rows)
    t0 = int(rows[i0]["Start_Timestamp"])
    lines = ["", f"Pass {pass_no}, dispatch by dispatch (times in us from the pass's first dispatch):", "",
             "| kernel | queue | start | end | us |", "|---|---|---|---|---|"]
    last = t0
    for r in rows[i0:i1]:
        a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        lines.append(f"| {short(r['Kernel_Name'])} | {r['Queue_Id']} | {(a - t0) / 1e3:.1f} | {(b - t0) / 1e3:.1f} | "
                     f"{(b - a) / 1e3:.1f} |")
        last = max(last, b)
    lines += ["", f"Pass {pass_no} spans {(last - t0) / 1e3:.1f} us."]
    return "\n".join(lines) + "\n"


def per_dispatch(csv_path):
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(csv_path)):
        agg[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return agg


def mfma_md(stats_csv, pmc_csv, cmd, peak_tflops=2500.0, simds=1024):
    """Per denoiser kernel: average duration (kernel trace), BF16 MFMA FLOPs
    per dispatch (SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512), the FLOP rate they
    imply, and MFMA busy fraction (SQ_VALU_MFMA_BUSY_CYCLES over
    GRBM_GUI_ACTIVE per XCD x SIMDs)."""
    dur = {}
    for r in csv.DictReader(open(stats_csv)):
        dur[short(r["Name"])] = (int(r["Calls"]), float(r["AverageNs"]))
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    xcds = 1
    for r in csv.DictReader(open(pmc_csv)):
        agg[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    lines = [f"Commands: {cmd}", "", "| kernel | calls | avg us | MFMA GFLOP / dispatch | TFLOP/s | MFMA busy % |",
             "|---|---|---|---|---|---|"]
    tot_ns = tot_flop = 0.0
    for k, (calls, ns) in sorted(dur.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        if "hh::dn::" not in k:
            continue
        c = agg.get(k, {})
        mops = sum(c.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", [0])) / max(len(c.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", [1])), 1)
        busy = sum(c.get("SQ_VALU_MFMA_BUSY_CYCLES", [0])) / max(len(c.get("SQ_VALU_MFMA_BUSY_CYCLES", [1])), 1)
        gui = sum(c.get("GRBM_GUI_ACTIVE", [0])) / max(len(c.get("GRBM_GUI_ACTIVE", [1])), 1)
        flop = mops * 512
        # GRBM_GUI_ACTIVE is summed over the 8 XCDs
        util = 100.0 * busy / (gui / 8 * simds) if gui else 0.0
        lines.append(f"| {k} | {calls} | {ns / 1e3:.1f} | {flop / 1e9:.2f} | {flop / ns / 1e3:.1f} | {util:.1f} |")
        if k not in ("hh::dn::k_relayout", "hh::dn::k_zero_f32", "hh::dn::k_to_bf16"):
            tot_ns += ns * calls
            tot_flop += flop * calls
    lines += ["", f"Forward passes: {tot_flop / 1e9:.1f} GFLOP of MFMA work in {tot_ns / 1e6:.3f} ms of kernel time "
              f"= {tot_flop / tot_ns / 1e3:.1f} TFLOP/s = {tot_flop / tot_ns / 1e3 / peak_tflops:.3f} of the "
              f"{peak_tflops:.0f} TFLOP/s dense BF16 peak."]
    return "\n".join(lines) + "\n"


def l2_traffic(pmc_csv, kernel, avg_ms, cmd):
    """L1 -> L2 read traffic of a kernel: TCP_TCC_READ_REQ per dispatch, in
    bytes by what the same counter reads for the 1 GiB copy kernel (128 B per
    request on gfx950), plus the L2 hit rate (TCC_HIT / (TCC_HIT + TCC_MISS))
    and the mean L1 -> L2 read latency in cycles."""
    c = per_dispatch_multi(pmc_csv)
    probe = c["k_copy_probe"]
    req_bytes = PROBE_BYTES / (sum(probe["TCP_TCC_READ_REQ_sum"]) / len(probe["TCP_TCC_READ_REQ_sum"]))
    k = [v for n, v in c.items() if kernel in n and "true" not in n][0]
    avg = {name: sum(vs) / len(vs) for name, vs in k.items()}
    rd = avg["TCP_TCC_READ_REQ_sum"] * req_bytes
    return {
        "command": cmd, "bytes_per_request": round(req_bytes, 1),
        "l1_accesses_per_launch": round(avg["TCP_TOTAL_CACHE_ACCESSES_sum"]),
        "l2_read_bytes_per_launch": round(rd),
        "l2_read_gbps": round(rd / (avg_ms * 1e-3) / 1e9, 1),
        "l1_hit_rate": round(1.0 - avg["TCP_TCC_READ_REQ_sum"] / avg["TCP_TOTAL_CACHE_ACCESSES_sum"], 4),
        "l2_hit_rate": round(avg["TCC_HIT_sum"] / (avg["TCC_HIT_sum"] + avg["TCC_MISS_sum"]), 4),
        "l1_to_l2_read_latency_cycles": round(avg["TCP_TCC_READ_REQ_LATENCY_sum"] / avg["TCP_TCC_READ_REQ_sum"], 1),
    }


def per_dispatch_multi(csv_path):
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(csv_path)):
        n = short(r["Kernel_Name"])
        agg[n.split("<")[0].split("::")[-1] if "k_copy_probe" in n else n][r["Counter_Name"]].append(
            float(r["Counter_Value"]))
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


ISSUE_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_BUSY_CYCLES",
                  "GRBM_GUI_ACTIVE")


def issue(pmc_csv, cmd, simds=1024, xcds=8, cycles_per_valu=4):
    """Issue-side limits of the traversal and shading kernels, per dispatch:

      valu_issue_frac  = SQ_INSTS_VALU x 4 cycles (a wave64 VALU instruction
                         occupies its SIMD for 4 cycles)
                         / (SIMDs x GRBM_GUI_ACTIVE / XCDs): the share of the
                         chip's VALU issue cycles the kernel's instructions use
      lane_utilization = SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU):
                         active lanes per issued VALU instruction
      valu_insts_per_wave = SQ_INSTS_VALU / SQ_WAVES
    GRBM_GUI_ACTIVE is summed over the XCDs."""
    agg = per_dispatch_multi(pmc_csv)
    out = {}
    for name, c in agg.items():
        base = name.split("<")[0].split("::")[-1]
        if base not in ("k_extend", "k_extend_packet", "k_shadow", "k_shade", "k_megakernel") or "true" in name:
            continue
        if any(k not in c for k in ISSUE_COUNTERS):
            continue
        avg = {k: sum(v) / len(v) for k, v in c.items()}
        cyc = avg["GRBM_GUI_ACTIVE"] / xcds
        ent = {
            "dispatches": len(c["SQ_WAVES"]),
            "valu_issue_frac": round(avg["SQ_INSTS_VALU"] * cycles_per_valu / (simds * cyc), 4) if cyc else None,
            "lane_utilization": round(avg["SQ_THREAD_CYCLES_VALU"] / (64 * avg["SQ_ACTIVE_INST_VALU"]), 4)
            if avg["SQ_ACTIVE_INST_VALU"] else None,
            "valu_insts_per_wave": round(avg["SQ_INSTS_VALU"] / avg["SQ_WAVES"], 1) if avg["SQ_WAVES"] else None,
            "raw": {k: round(avg[k], 1) for k in ISSUE_COUNTERS},
        }
        prev = out.get(base)
        if prev is None or prev["raw"]["GRBM_GUI_ACTIVE"] < ent["raw"]["GRBM_GUI_ACTIVE"]:
            out[base] = ent   # the instantiation that ran longest (the timed one)
    return {"command": cmd, "formula": " ".join(issue.__doc__.split()), "kernels": out}


def main():
    p = argparse.ArgumentParser()
    sub = p.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--cmd", required=True)
    s.add_argument("--title", default="")
    s.add_argument("--trace", default=None, help="kernel_trace.csv of the same run: per-pass extension launches")
    s.add_argument("--note", default="")
    s.add_argument("--timeline", type=int, default=0, help="with --trace: list every dispatch of this pass")
    s.add_argument("--out", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--kernel", default="k_extend")
    t.add_argument("--cmd", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--l2-csv", default=None, help="counter run with TCP_TCC_READ_REQ_sum, TCC_HIT_sum, ...")
    t.add_argument("--l2-cmd", default="")
    t.add_argument("--avg-ms", type=float, default=0.0, help="the kernel's average launch time")
    t.add_argument("--shape", default=None, help="JSON launch shape the counters were taken at")
    i = sub.add_parser("issue")
    i.add_argument("pmc_csv")
    i.add_argument("--cmd", required=True)
    i.add_argument("--shape", default=None)
    i.add_argument("--out", required=True)
    m = sub.add_parser("mfma")
    m.add_argument("stats_csv")
    m.add_argument("pmc_csv")
    m.add_argument("--cmd", required=True)
    m.add_argument("--out", required=True)
    a = p.parse_args()
    if a.what == "stats":
        text = (f"# {a.title}\n\n" if a.title else "") + stats_md(a.csv, a.cmd)
        if a.trace:
            text += passes_md(a.trace)
            if a.timeline:
                text += timeline_md(a.trace, a.timeline)
        if a.note:
            text += "\n" + a.note + "\n"
        Path(a.out).write_text(text)
    elif a.what == "mfma":
        Path(a.out).write_text(mfma_md(a.stats_csv, a.pmc_csv, a.cmd))
    elif a.what == "issue":
        res = issue(a.pmc_csv, a.cmd)
        if a.shape:
            res["shape"] = json.loads(a.shape)
        Path(a.out).write_text(json.dumps(res, indent=1) + "\n")
    else:
        res = traffic(a.fetch_csv, a.write_csv, a.kernel, a.cmd)
        if a.l2_csv:
            res["l2"] = l2_traffic(a.l2_csv, a.kernel, a.avg_ms, a.l2_cmd)
        if a.shape:
            res["shape"] = json.loads(a.shape)
        Path(a.out).write_text(json.dumps(res, indent=1) + "\n")
    print("wrote", a.out)


if __name__ == "__main__":
    main()
