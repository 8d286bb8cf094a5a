#!/bin/bash
# Profiling recipe behind profiles/*soft_block_21* (CELLS=20, default) and
# profiles/*soft_block_56* (CELLS=55, the 1M-tet block): kernel trace of the
# benchmark, then one rocprofv3 PMC pass per counter (FETCH_SIZE, WRITE_SIZE)
# on the benchmark and on the copy calibration kernel, then one pass of wave,
# occupancy and LDS counters.  Run from the repo
# root on the GPU host; writes raw output under gpurun_out/prof_* and the
# summaries via summarize.py.
set -u
cd "$(dirname "$0")/../.."
export TMPDIR=/tmp
OUT=gpurun_out
CELLS=${CELLS:-20}
BENCH="bench.py --cells $CELLS --steps 4 --warmup 2 --no-cpu --no-learned --quick --res-frames 0 --fp64-steps 0 --lanes ${LANES:-32} --sweep ${SWEEP:-1}"
/opt/rocm/bin/hipcc -O2 --offload-arch=gfx950 -o avbd-metal_amd/lib/copy_calib avbd-metal_amd/prof/copy_calib.hip || exit 1
echo "pass 1: kernel trace"
timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/prof_trace -o trace -- python3 $BENCH > $OUT/prof_trace.log 2>&1
rc=$?; tail -2 $OUT/prof_trace.log; [ -f $OUT/prof_trace/trace_kernel_stats.csv ] || exit $rc
for C in FETCH_SIZE WRITE_SIZE; do
  echo "pass: $C bench"
  timeout -s KILL 120 rocprofv3 --pmc $C --output-format csv -d $OUT/prof_$C -o bench -- python3 $BENCH > $OUT/prof_$C.log 2>&1
  rc=$?; tail -1 $OUT/prof_$C.log; [ -f $OUT/prof_$C/bench_counter_collection.csv ] || exit $rc
  echo "pass: $C copy calibration"
  timeout -s KILL 60 rocprofv3 --pmc $C --output-format csv -d $OUT/prof_$C -o copy -- ./avbd-metal_amd/lib/copy_calib > $OUT/prof_${C}_copy.log 2>&1
  rc=$?; tail -1 $OUT/prof_${C}_copy.log; [ -f $OUT/prof_$C/copy_counter_collection.csv ] || exit $rc
done
# cross-check of the byte counters: the L2 -> fabric read requests by size
# class (128 B requests count in TCC_BUBBLE) and the DRAM-bound ones
XC="TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_BUBBLE_sum TCC_EA0_RDREQ_DRAM_sum"
echo "pass: read-request counters bench"
timeout -s KILL 120 rocprofv3 --pmc $XC --output-format csv -d $OUT/prof_XC -o bench -- python3 $BENCH > $OUT/prof_XC.log 2>&1
rc=$?; tail -1 $OUT/prof_XC.log; [ -f $OUT/prof_XC/bench_counter_collection.csv ] || exit $rc
echo "pass: read-request counters copy calibration"
timeout -s KILL 60 rocprofv3 --pmc $XC --output-format csv -d $OUT/prof_XC -o copy -- ./avbd-metal_amd/lib/copy_calib > $OUT/prof_XC_copy.log 2>&1
rc=$?; tail -1 $OUT/prof_XC_copy.log; [ -f $OUT/prof_XC/copy_counter_collection.csv ] || exit $rc
echo "pass: wave / LDS counters"
timeout -s KILL 120 rocprofv3 --pmc SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT GRBM_GUI_ACTIVE \
  --output-format csv -d $OUT/prof_SQ -o bench -- python3 $BENCH > $OUT/prof_SQ.log 2>&1
rc=$?; tail -1 $OUT/prof_SQ.log; [ -f $OUT/prof_SQ/bench_counter_collection.csv ] || exit $rc
python3 avbd-metal_amd/prof/summarize.py "$OUT" "$BENCH" $((CELLS + 1))
