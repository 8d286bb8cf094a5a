#!/bin/bash
# Profiling recipe behind profiles/kernels_batched_4096.txt: kernel trace of
# the 4096 batched rigid scenes benchmark (94001 bodies, contact-aware
# dynamic coloring), summarized per kernel.  Run from the repo root on the
# GPU host; raw output under gpurun_out/prof_batched.
set -u
cd "$(dirname "$0")/../.."
export TMPDIR=/tmp
OUT=gpurun_out
BENCH="bench.py --workload batched --steps 4 --warmup 4 --no-cpu --lanes ${LANES:-4} --sweep ${SWEEP:-1}"
timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/prof_batched -o trace -- python3 $BENCH > $OUT/prof_batched.log 2>&1
rc=$?; tail -2 $OUT/prof_batched.log; [ -f $OUT/prof_batched/trace_kernel_stats.csv ] || exit $rc
python3 avbd-metal_amd/prof/summarize.py --kernels-only "$OUT/prof_batched/trace_kernel_stats.csv" \
  "$OUT/summary/kernels_batched_4096.txt" \
  "rocprofv3 kernel trace, 4096 batched rigid scenes (94001 bodies, 25348 hinges), 8 frames, 10 iterations" \
  "avbd-metal_amd/prof/profile_batched.sh  (rocprofv3 --kernel-trace --stats --output-format csv -- python3 $BENCH)"
