#!/bin/bash
# Profiling recipe behind profiles/kernels_cloth_256.txt: kernel trace of the
# 256^2 cloth drape benchmark (a few frames after contact), summarized per
# kernel.  Run from the repo root on the GPU host; raw output under
# gpurun_out/prof_cloth.
set -u
cd "$(dirname "$0")/../.."
export TMPDIR=/tmp
OUT=gpurun_out
BENCH="bench.py --workload cloth --steps 4 --warmup 40 --no-cpu --lanes ${LANES:-8} --sweep ${SWEEP:-1}"
timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/prof_cloth -o trace -- python3 $BENCH > $OUT/prof_cloth.log 2>&1
rc=$?; tail -2 $OUT/prof_cloth.log; [ -f $OUT/prof_cloth/trace_kernel_stats.csv ] || exit $rc
python3 avbd-metal_amd/prof/summarize.py --kernels-only "$OUT/prof_cloth/trace_kernel_stats.csv" \
  "$OUT/summary/kernels_cloth_256.txt" \
  "rocprofv3 kernel trace, cloth drape 256^2 (65536 vertices, 130050 triangles), 44 frames, 30 iterations" \
  "avbd-metal_amd/prof/profile_cloth.sh  (rocprofv3 --kernel-trace --stats --output-format csv -- python3 $BENCH)"
