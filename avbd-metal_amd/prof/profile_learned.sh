#!/bin/bash
# Profiling recipe behind profiles/kernels_learned_soft_block_21.txt: kernel
# trace of the learned warm start's kernels -- GPU training on synthetic rows,
# then fixed-iteration frames of the 21^3 block with and without the learned
# start (avbd-metal_amd/prof/learned_frames.py) -- summarized per kernel.
# Run from the repo root on the GPU host; raw output under
# gpurun_out/prof_learned.
set -u
cd "$(dirname "$0")/../.."
export TMPDIR=/tmp
OUT=gpurun_out
RUN="avbd-metal_amd/prof/learned_frames.py"
timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/prof_learned -o trace -- python3 -u $RUN > $OUT/prof_learned.log 2>&1
rc=$?; tail -4 $OUT/prof_learned.log; [ -f $OUT/prof_learned/trace_kernel_stats.csv ] || exit $rc
python3 avbd-metal_amd/prof/summarize.py --kernels-only "$OUT/prof_learned/trace_kernel_stats.csv" \
  "$OUT/summary/kernels_learned_soft_block_21.txt" \
  "rocprofv3 kernel trace, learned warm start: GPU training on 65536 synthetic rows, 20 plain + 20 learned frames of the soft block 21^3" \
  "avbd-metal_amd/prof/profile_learned.sh  (rocprofv3 --kernel-trace --stats --output-format csv -- python3 -u $RUN)"
