#!/bin/bash
# Profiling recipe behind profiles/kernels_learned_soft_block_21.txt: kernel
# trace of a shortened learned-warm-start comparison on the 21^3 block
# (sample collection, GPU training, frames with the learned stage), summarized
# per kernel.  Run from the repo root on the GPU host; raw output under
# gpurun_out/prof_learned.
set -u
cd "$(dirname "$0")/../.."
export TMPDIR=/tmp
OUT=gpurun_out
BENCH="bench.py --steps 4 --warmup 2 --no-cpu --quick --fp64-steps 0 --res-frames 6 --res-cap 300 --train-seeds 1 --train-cloth-seeds 1 --train-frames 10 --label-cap 300 --epochs 4 --lanes ${LANES:-32} --sweep ${SWEEP:-1}"
timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/prof_learned -o trace -- python3 $BENCH > $OUT/prof_learned.log 2>&1
rc=$?; tail -2 $OUT/prof_learned.log; [ -f $OUT/prof_learned/trace_kernel_stats.csv ] || exit $rc
python3 avbd-metal_amd/prof/summarize.py --kernels-only "$OUT/prof_learned/trace_kernel_stats.csv" \
  "$OUT/summary/kernels_learned_soft_block_21.txt" \
  "rocprofv3 kernel trace, learned warm start on the soft block 21^3 (collection, training, learned frames)" \
  "avbd-metal_amd/prof/profile_learned.sh  (rocprofv3 --kernel-trace --stats --output-format csv -- python3 $BENCH)"
