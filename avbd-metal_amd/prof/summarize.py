"""Summaries of the profiling recipe (profile_soft_block.sh): per-kernel
durations of the benchmark and HBM bytes per launch of the primal kernel
from the FETCH_SIZE / WRITE_SIZE passes, calibrated on a plain copy of known
size.  Writes kernels_soft_block_21.txt and primal_soft_block_21_pmc.json
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


def main():
    out, bench_cmd = sys.argv[1], sys.argv[2]
    prof = os.path.join(out, "summary")
    os.makedirs(prof, exist_ok=True)
    stats = kernel_stats(os.path.join(out, "prof_trace", "trace_kernel_stats.csv"))
    cmd = ("avbd-metal_amd/prof/profile_soft_block.sh  (rocprofv3 --kernel-trace --stats --output-format csv "
           f"-- python3 {bench_cmd})")
    lines = [f"# rocprofv3 kernel trace, soft block 21^3 (9261 vertices, 48000 tets, 20 iterations)",
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
    if res["FETCH_SIZE"]["primal_bytes"] is not None and res["WRITE_SIZE"]["primal_bytes"] is not None:
        res["hbm_bytes_per_launch"] = res["FETCH_SIZE"]["primal_bytes"] + res["WRITE_SIZE"]["primal_bytes"]
        lines.append(f"# HBM bytes per k_primal launch (calibrated): {res['hbm_bytes_per_launch']:.0f}")
    with open(os.path.join(prof, "kernels_soft_block_21.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(prof, "primal_soft_block_21_pmc.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("\n".join(lines[-4:]))


if __name__ == "__main__":
    main()
