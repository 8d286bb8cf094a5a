"""Summaries of the profiling recipes (profile_soft_block.sh,
profile_learned.sh): per-kernel
durations of the benchmark and HBM bytes per launch of the primal kernel
from the FETCH_SIZE / WRITE_SIZE passes, calibrated on a plain copy of known
size.  Writes kernels_soft_block_<side>.txt and primal_soft_block_<side>_pmc.json
into <out>/summary (copied into profiles/ afterwards)."""
import csv
import json
import os
import sys
from collections import defaultdict

COPY_BYTES = 256 << 20


def kernel_stats(path):
    rows = list(csv.DictReader(open(path)))
    return sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))


def counter_avg(path, counter, name_filter):
    vals = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name") != counter:
            continue
        if name_filter not in r.get("Kernel_Name", ""):
            continue
        vals[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    allv = [v for vs in vals.values() for v in vs]
    return (sum(allv) / len(allv) if allv else None), len(allv)


def kernels_only(stats_csv, dest, title, cmd):
    """Per-kernel table of one kernel trace (no counters)."""
    stats = kernel_stats(stats_csv)
    lines = [f"# {title}", f"# command: {cmd}", "#",
             f"# {'kernel':70s} {'calls':>6s} {'avg_us':>9s} {'total_ms':>9s} {'pct':>6s}"]
    for r in stats[:40]:
        lines.append(f"{r['Name'][:72]:72s} {r['Calls']:>6s} {float(r['AverageNs']) / 1e3:9.3f} "
                     f"{float(r['TotalDurationNs']) / 1e6:9.3f} {float(r['Percentage']):6.2f}")
    os.makedirs(os.path.dirname(dest), exist_ok=True)
    with open(dest, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines[:12]))


def main():
    if sys.argv[1] == "--kernels-only":
        kernels_only(*sys.argv[2:6])
        return
    out, bench_cmd = sys.argv[1], sys.argv[2]
    side = int(sys.argv[3]) if len(sys.argv) > 3 else 21
    prof = os.path.join(out, "summary")
    os.makedirs(prof, exist_ok=True)
    stats = kernel_stats(os.path.join(out, "prof_trace", "trace_kernel_stats.csv"))
    cmd = (f"CELLS={side - 1} avbd-metal_amd/prof/profile_soft_block.sh  (rocprofv3 --kernel-trace --stats --output-format csv "
           f"-- python3 {bench_cmd})")
    cells = side - 1
    lines = [f"# rocprofv3 kernel trace, soft block {side}^3 ({side ** 3} vertices, {6 * cells ** 3} tets, 20 iterations)",
             f"# command: {cmd}", "#",
             f"# {'kernel':70s} {'calls':>6s} {'avg_us':>9s} {'total_ms':>9s} {'pct':>6s}"]
    for r in stats[:30]:
        lines.append(f"{r['Name'][:72]:72s} {r['Calls']:>6s} {float(r['AverageNs']) / 1e3:9.3f} "
                     f"{float(r['TotalDurationNs']) / 1e6:9.3f} {float(r['Percentage']):6.2f}")
    primal = [r for r in stats if "k_primal" in r["Name"]]
    res = {"command": cmd}
    if primal:
        res["primal_kernel"] = primal[0]["Name"]
        res["primal_avg_us"] = float(primal[0]["AverageNs"]) / 1e3
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        b, nb = counter_avg(os.path.join(out, f"prof_{c}", "bench_counter_collection.csv"), c, "k_primal")
        k, nk = counter_avg(os.path.join(out, f"prof_{c}", "copy_counter_collection.csv"), c, "k_copy")
        factor = COPY_BYTES / (k * 1024.0) if k else None
        res[c] = {"primal_avg_kib": b, "primal_dispatches": nb, "copy_avg_kib": k, "copy_bytes": COPY_BYTES,
                  "bytes_per_kib": factor * 1024.0 if factor else None,
                  "primal_bytes": b * 1024.0 * factor if (b is not None and factor) else None}
        lines.append(f"# {c}: k_primal avg {b} KiB over {nb} dispatches; copy kernel {k} KiB for {COPY_BYTES} "
                     f"bytes -> factor {factor}")
    xc_dir = os.path.join(out, "prof_XC")
    if os.path.exists(os.path.join(xc_dir, "bench_counter_collection.csv")):
        # read requests by size class: bytes = 32 * RDREQ_32B + 128 * BUBBLE
        # + 64 * (the rest), the expression FETCH_SIZE itself uses; the DRAM
        # request count beside it, per k_primal launch and for the copy
        xc = {}
        for tag, flt in (("bench", "k_primal"), ("copy", "k_copy")):
            path = os.path.join(xc_dir, f"{tag}_counter_collection.csv")
            v = {c: counter_avg(path, c, flt)[0] for c in ("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum",
                                                        "TCC_BUBBLE_sum", "TCC_EA0_RDREQ_DRAM_sum")}
            if None not in v.values():
                rest = v["TCC_EA0_RDREQ_sum"] - v["TCC_BUBBLE_sum"] - v["TCC_EA0_RDREQ_32B_sum"]
                v["bytes_by_size_class"] = 32 * v["TCC_EA0_RDREQ_32B_sum"] + 128 * v["TCC_BUBBLE_sum"] + 64 * rest
                v["dram_requests_x128"] = 128 * v["TCC_EA0_RDREQ_DRAM_sum"]
            xc["primal" if tag == "bench" else "copy"] = v
        res["read_request_cross_check"] = xc
        cp = xc.get("copy", {})
        if cp.get("bytes_by_size_class"):
            lines.append(f"# read requests, copy of {COPY_BYTES} B: {cp['bytes_by_size_class']:.0f} B by size class, "
                         f"{cp['dram_requests_x128']:.0f} B as 128 B DRAM requests")
        pr = xc.get("primal", {})
        if pr.get("bytes_by_size_class"):
            lines.append(f"# read requests, k_primal per launch: {pr['bytes_by_size_class']:.0f} B by size class, "
                         f"{pr['dram_requests_x128']:.0f} B as 128 B DRAM requests")
    sq_path = os.path.join(out, "prof_SQ", "bench_counter_collection.csv")
    if os.path.exists(sq_path):
        sq = {}
        for c in ("SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT",
                  "GRBM_GUI_ACTIVE"):
            v, nv = counter_avg(sq_path, c, "k_primal")
            sq[c] = v
        res["primal_wave_counters_avg_per_dispatch"] = sq
        if sq.get("SQ_WAVE_CYCLES") and sq.get("SQ_BUSY_CYCLES"):
            # waves resident per busy SQ cycle, summed over the SQs that count them
            res["primal_waves_per_busy_cycle"] = sq["SQ_WAVE_CYCLES"] / sq["SQ_BUSY_CYCLES"]
        lines.append("# k_primal per dispatch: " + ", ".join(f"{k} {v:.0f}" for k, v in sq.items() if v is not None))
    if res["FETCH_SIZE"]["primal_bytes"] is not None and res["WRITE_SIZE"]["primal_bytes"] is not None:
        res["hbm_bytes_per_launch"] = res["FETCH_SIZE"]["primal_bytes"] + res["WRITE_SIZE"]["primal_bytes"]
        lines.append(f"# HBM bytes per k_primal launch (calibrated): {res['hbm_bytes_per_launch']:.0f}")
    with open(os.path.join(prof, f"kernels_soft_block_{side}.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(prof, f"primal_soft_block_{side}_pmc.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("\n".join(lines[-4:]))


if __name__ == "__main__":
    main()
