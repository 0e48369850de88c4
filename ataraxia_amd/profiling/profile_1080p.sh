#!/bin/bash
# rocprofv3 passes behind profiles/plain_1080p_k_render_<kernel>_<accel>.txt.
# Run from the repo root on a machine with the GPU:
#     KERNEL=plain|wavefront ACCEL=bvh|none [OUT=dir] bash ataraxia_amd/profiling/profile_1080p.sh
# then summarise with: python ataraxia_amd/profiling/summarize.py <OUT>
#   1. kernel trace + stats of bench.py (STEPS frames of the 1080p workload;
#      wavefront: 256 frames, the renderer's launch length, so one launch)
#   2. FETCH_SIZE, then WRITE_SIZE, each in its own pass, over one launch:
#      PMC_STEPS frames (plain: one launch per frame; wavefront: 256 frames
#      per launch, so PMC_STEPS=256 is exactly one launch)
#   3. the same two counters on copy_probe (256 MiB copy) for calibration
#   4. SQ counters (waves, VALU instructions and lane utilisation, waits, loads)
# Every pass has its own time limit; the script stops at the first failure.
set -u
REPO=$(pwd)
KERNEL=${KERNEL:-wavefront}
ACCEL=${ACCEL:-bvh}
OUT=${OUT:-$REPO/profile_out/prof_${KERNEL}_$ACCEL}
case "$OUT" in /*) ;; *) OUT="$REPO/$OUT" ;; esac   # the passes run from /tmp
if [ "$KERNEL" = plain ]; then STEPS=${STEPS:-128}; PMC_STEPS=${PMC_STEPS:-4}; else STEPS=${STEPS:-256}; PMC_STEPS=${PMC_STEPS:-256}; fi
BENCH=("$REPO/bench.py" --kernel "$KERNEL" --accel "$ACCEL" --no-serve --no-cpu-baseline)
mkdir -p "$OUT"
export TMPDIR=/tmp
run() { echo "== $*"; "$@"; local rc=$?; echo "== rc=$rc"; return $rc; }
cd /tmp || exit 1
run timeout -k 10 240 rocprofv3 --kernel-trace --stats -d "$OUT/kt" -o k -- \
    python3 "${BENCH[@]}" --steps "$STEPS" --warmup 0 &&
run timeout -s KILL 90 rocprofv3 --pmc FETCH_SIZE -d "$OUT/fetch" -o k -- \
    python3 "${BENCH[@]}" --steps "$PMC_STEPS" --warmup 0 &&
run timeout -s KILL 90 rocprofv3 --pmc WRITE_SIZE -d "$OUT/write" -o k -- \
    python3 "${BENCH[@]}" --steps "$PMC_STEPS" --warmup 0 &&
run timeout -s KILL 60 rocprofv3 --pmc FETCH_SIZE -d "$OUT/copy_fetch" -o copy -- \
    "$REPO/ataraxia_amd/profiling/copy_probe" 268435456 &&
run timeout -s KILL 60 rocprofv3 --pmc WRITE_SIZE -d "$OUT/copy_write" -o copy -- \
    "$REPO/ataraxia_amd/profiling/copy_probe" 268435456 &&
run timeout -s KILL 90 rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_WAIT_INST_ANY SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_VMEM_RD -d "$OUT/sq" -o k -- \
    python3 "${BENCH[@]}" --steps "$PMC_STEPS" --warmup 0
