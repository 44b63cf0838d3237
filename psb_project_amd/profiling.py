"""Profiling helpers: turn rocprofv3 output into the short text summaries kept
under profiles/, and a copy program of known size that calibrates the HBM
traffic counters.

    python -m psb_project_amd.profiling copy [--mib 512] [--reps 8]
        runs `reps` streaming copies of `mib` MiB through
        psb_copy_array_device_to_device (kernel k_copy_stream): each reads and
        writes exactly mib MiB, so FETCH_SIZE / WRITE_SIZE of that kernel give
        the counter-to-bytes factor;
    python -m psb_project_amd.profiling kernel-stats STATS_CSV --command CMD --out FILE
            [--trace KERNEL_TRACE_CSV] [--replay N]
        formats rocprofv3 --stats kernel_stats.csv (top kernels by time) and,
        from the per-dispatch trace, the average over full dispatches (and
        over the last N dispatches: the bench's corrector replay);
    python -m psb_project_amd.profiling counters CSV [CSV ...] --kernel REGEX --cus N
            --commands TEXT --out FILE
        per-kernel averages of every counter in rocprofv3 --pmc counter
        collections (dispatches of at least 5 us), per wavefront where that
        reads better, and the derived issue and busy fractions;
    python -m psb_project_amd.profiling traffic --kernel REGEX --fetch CSV --write CSV
            --copy-fetch CSV --copy-write CSV --copy-bytes B --alg-bytes A
            --commands TEXT --out FILE
        average FETCH_SIZE and WRITE_SIZE (KiB) per dispatch of the kernels
        matching REGEX, each scaled by known/measured of the copy kernel, and
        writes hbm_bytes_per_launch = fetch + write.
"""
from __future__ import annotations

import argparse
import csv
import re
import sys
from collections import defaultdict


def _rows(path):
    with open(path, newline="") as f:
        yield from csv.DictReader(f)


def counter_per_dispatch(path, kernel_regex: str, counter: str, min_us: float = 0.0):
    """Mean value of `counter` per dispatch of kernels matching the regex,
    over dispatches that ran at least min_us (drops the corrector's
    early-exit launches)."""
    rx = re.compile(kernel_regex)
    per = defaultdict(float)
    for r in _rows(path):
        if r.get("Counter_Name") != counter or not rx.search(r.get("Kernel_Name", "")):
            continue
        if min_us > 0 and "Start_Timestamp" in r:
            if (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 < min_us:
                continue
        per[r.get("Dispatch_Id") or r.get("Correlation_Id")] += float(r["Counter_Value"])
    if not per:
        raise SystemExit(f"{path}: no {counter} rows for kernels matching {kernel_regex!r}")
    return sum(per.values()) / len(per), len(per)


def cmd_copy(a) -> int:
    import numpy as np
    from ._lib import check, lib
    from .device import Context
    ctx = Context(a.device)
    nbytes = a.mib << 20
    src = ctx.array(nbytes)
    dst = ctx.array(nbytes)
    check(lib().psb_memset_array(ctx.handle, src.ptr, 1, nbytes))
    for _ in range(a.reps):
        check(lib().psb_copy_array_device_to_device(ctx.handle, dst.ptr, src.ptr, 0, nbytes))
    ctx.sync()
    probe = np.empty(4, np.uint8)
    check(lib().psb_copy_array_from_device(ctx.handle, probe.ctypes.data, dst.ptr, nbytes - 4, 4))
    assert np.all(probe == 1)
    print(f"copied {a.reps} x {nbytes} bytes")
    src.free()
    dst.free()
    ctx.close()
    return 0


def cmd_kernel_stats(a) -> int:
    rows = list(_rows(a.stats_csv))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = [f"# rocprofv3 kernel statistics: {a.title}", "",
           f"command: {a.command}", "",
           f"{'kernel':60s} {'calls':>7s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'share':>6s}"]
    for r in rows[:a.top]:
        name = r["Name"]
        short = re.sub(r"\(.*$", "", name).replace("void ", "")
        out.append(f"{short[:60]:60s} {int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:9.3f} "
                   f"{float(r['MinNs']) / 1e3:9.3f} {float(r['MaxNs']) / 1e3:9.3f} "
                   f"{100 * float(r['TotalDurationNs']) / total:5.1f}%")
    if a.trace:
        # Speculative corrector launches queued past the converged iteration
        # read one status word and return; split them off by duration so the
        # average over full dispatches can be set beside the benchmark's.
        per = defaultdict(list)
        rx = re.compile(a.split_regex)
        trace = sorted(_rows(a.trace), key=lambda r: int(r["Start_Timestamp"]))
        for r in trace:
            if not rx.search(r["Kernel_Name"]):
                continue
            name = re.sub(r"\(.*$", "", r["Kernel_Name"]).replace("void ", "")
            per[name].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        out += ["", f"full dispatches (>= {a.full_min_us} us) vs early exits, from the kernel trace:",
                f"{'kernel':60s} {'full':>7s} {'avg_us':>9s} {'early':>7s} {'early_us':>9s}"]
        for name, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
            full = [x for x in d if x >= a.full_min_us]
            early = [x for x in d if x < a.full_min_us]
            if not full:
                continue
            eavg = sum(early) / len(early) if early else 0.0
            out.append(f"{name[:60]:60s} {len(full):7d} {sum(full) / len(full):9.3f} "
                       f"{len(early):7d} {eavg:9.3f}")
            if len(out) > 60:
                break
        if a.replay:
            # bench.py ends its timed run by replaying the corrector passes
            # back to back on the final state (its roofline): the last
            # `replay` dispatches of each split kernel
            out += ["", f"corrector passes replayed on the final state (the last {a.replay} "
                        f"dispatches of each):",
                    f"{'kernel':60s} {'replay':>7s} {'avg_us':>9s}"]
            for name, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
                last = d[-a.replay:]
                out.append(f"{name[:60]:60s} {len(last):7d} {sum(last) / len(last):9.3f}")
    out.append("")
    text = "\n".join(out)
    with open(a.out, "w") as f:
        f.write(text)
    print(text)
    return 0


def cmd_counters(a) -> int:
    """Average every counter per dispatch, per kernel, over the given
    counter-collection files (one per --pmc pass), plus derived figures:
    VALU issue time at 4 cycles per wave64 instruction on 4 SIMDs per CU,
    TA/TD busy per CU, L1 lines per vector read -- as fractions of the
    dispatch duration at the given clock."""
    rx = re.compile(a.kernel)
    vals = defaultdict(lambda: defaultdict(list))
    durs = defaultdict(list)
    for path in a.csv:
        per = defaultdict(lambda: defaultdict(float))
        meta = {}
        for r in _rows(path):
            k = r.get("Kernel_Name", "")
            if not rx.search(k):
                continue
            d = r.get("Dispatch_Id") or r.get("Correlation_Id")
            per[d][r["Counter_Name"]] += float(r["Counter_Value"])
            meta[d] = (re.sub(r"\(.*$", "", k).replace("void ", ""),
                       (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        for d, cs in per.items():
            name, us = meta[d]
            if us < a.full_min_us:
                continue
            durs[(name, path)].append(us)
            for c, v in cs.items():
                vals[name][c].append(v)
    out = [f"# rocprofv3 counters: {a.title}", "", "commands:",
           *[f"  {c}" for c in a.commands.split(";;")], "",
           f"dispatches of at least {a.full_min_us} us; {a.cus} CUs x 4 SIMDs, {a.ghz} GHz", ""]
    for name in sorted(vals):
        cs = {c: sum(v) / len(v) for c, v in vals[name].items()}
        us = [x for (n, _), l in durs.items() if n == name for x in l]
        dur = sum(us) / len(us)
        cyc = dur * 1e3 * a.ghz
        out.append(f"{name}  (mean duration under counters {dur:.2f} us)")
        waves = cs.get("SQ_WAVES")
        for c in sorted(cs):
            line = f"  {c:32s} {cs[c]:16.1f}"
            if waves and c.startswith("SQ_INSTS"):
                line += f"   {cs[c] / waves:10.1f} per wavefront"
            out.append(line)
        if "SQ_INSTS_VALU" in cs:
            issue = cs["SQ_INSTS_VALU"] / (4 * a.cus) * 4
            out.append(f"  -> VALU issue {issue / (a.ghz * 1e3):.2f} us = {issue / cyc:.0%} of the dispatch")
        for c, lab in (("TA_TA_BUSY_sum", "TA busy"), ("TD_TD_BUSY_sum", "TD busy")):
            if c in cs:
                out.append(f"  -> {lab} {cs[c] / a.cus / cyc:.0%} per CU")
        if "TCP_TOTAL_CACHE_ACCESSES_sum" in cs and cs.get("SQ_INSTS_VMEM_RD"):
            out.append(f"  -> L1 accesses per vector read "
                       f"{cs['TCP_TOTAL_CACHE_ACCESSES_sum'] / cs['SQ_INSTS_VMEM_RD']:.1f}")
        out.append("")
    text = "\n".join(out)
    with open(a.out, "w") as f:
        f.write(text)
    print(text)
    return 0


def cmd_traffic(a) -> int:
    fk, nd = counter_per_dispatch(a.fetch, a.kernel, "FETCH_SIZE", a.full_min_us)
    wk, _ = counter_per_dispatch(a.write, a.kernel, "WRITE_SIZE", a.full_min_us)
    cf, _ = counter_per_dispatch(a.copy_fetch, "k_copy_stream", "FETCH_SIZE")
    cw, _ = counter_per_dispatch(a.copy_write, "k_copy_stream", "WRITE_SIZE")
    ff = a.copy_bytes / (cf * 1024.0)
    fw = a.copy_bytes / (cw * 1024.0)
    fetch_b = fk * 1024.0 * ff
    write_b = wk * 1024.0 * fw
    hbm = fetch_b + write_b
    lines = [
        f"# HBM traffic per launch: {a.title}", "",
        "commands:", *[f"  {c}" for c in a.commands.split(";;")], "",
        f"kernel regex: {a.kernel}   dispatches averaged: {nd} (those >= {a.full_min_us} us)",
        f"FETCH_SIZE per dispatch: {fk:.1f} KiB (raw)",
        f"WRITE_SIZE per dispatch: {wk:.1f} KiB (raw)",
        f"copy calibration (k_copy_stream, {a.copy_bytes} B read and written per dispatch):",
        f"  FETCH_SIZE {cf:.1f} KiB -> factor {ff:.4f};  WRITE_SIZE {cw:.1f} KiB -> factor {fw:.4f}",
        f"fetch_bytes_per_launch = {fetch_b:.0f}",
        f"write_bytes_per_launch = {write_b:.0f}",
        f"hbm_bytes_per_launch = {hbm:.0f}",
        f"alg_bytes_per_launch = {a.alg_bytes:.0f}",
        f"traffic / algorithmic = {hbm / a.alg_bytes:.3f}",
        "",
    ]
    text = "\n".join(lines)
    with open(a.out, "w") as f:
        f.write(text)
    print(text)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=512)
    c.add_argument("--reps", type=int, default=8)
    c.add_argument("--device", type=int, default=0)
    k = sub.add_parser("kernel-stats")
    k.add_argument("stats_csv")
    k.add_argument("--command", required=True)
    k.add_argument("--title", default="")
    k.add_argument("--top", type=int, default=20)
    k.add_argument("--trace", default=None, help="kernel_trace.csv of the same run")
    k.add_argument("--full-min-us", type=float, default=5.0)
    k.add_argument("--split-regex", default=r"k_pci_(density|force)",
                   help="kernels whose early-exit launches are split off")
    k.add_argument("--replay", type=int, default=0,
                   help="average the last N dispatches of each split kernel separately")
    k.add_argument("--out", required=True)
    q = sub.add_parser("counters")
    q.add_argument("csv", nargs="+")
    q.add_argument("--kernel", required=True)
    q.add_argument("--cus", type=int, default=256)
    q.add_argument("--ghz", type=float, default=2.4)
    q.add_argument("--commands", required=True)
    q.add_argument("--title", default="")
    q.add_argument("--full-min-us", type=float, default=5.0)
    q.add_argument("--out", required=True)
    t = sub.add_parser("traffic")
    for name in ("kernel", "fetch", "write", "copy_fetch", "copy_write", "commands", "out"):
        t.add_argument("--" + name.replace("_", "-"), required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--alg-bytes", type=float, required=True)
    t.add_argument("--title", default="")
    t.add_argument("--full-min-us", type=float, default=5.0)
    a = ap.parse_args(argv)
    return {"copy": cmd_copy, "kernel-stats": cmd_kernel_stats, "traffic": cmd_traffic,
            "counters": cmd_counters}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
