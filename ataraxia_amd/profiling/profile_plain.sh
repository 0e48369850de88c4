#!/bin/bash
# rocprofv3 passes behind profiles/plain_1080p_*.txt.  Run from the repo root
# on the GPU host:  bash ataraxia_amd/profiling/profile_plain.sh
#   1. kernel trace + stats of bench.py (16 frames of the 1080p workload)
#   2. FETCH_SIZE, then WRITE_SIZE of 4 frames, each in its own pass
#   3. the same two counters on copy_probe (256 MiB copy) for calibration
#   4. SQ counters (waves, VALU instructions, active VALU cycles, LDS)
# Every pass has its own time limit; the script stops at the first failure.
set -u
REPO=$(pwd)
OUT=${OUT:-$REPO/gpurun_out/prof}
STEPS=${STEPS:-16}
mkdir -p "$OUT"
export TMPDIR=/tmp
run() { echo "== $*"; "$@"; local rc=$?; echo "== rc=$rc"; return $rc; }
cd /tmp || exit 1
timeout -k 5 60 rocprofv3 --list-avail > "$OUT/avail.txt" 2>&1 || true
run timeout -k 10 240 rocprofv3 --kernel-trace --stats -d "$OUT/kt" -o plain -- \
    python3 "$REPO/bench.py" --steps "$STEPS" --warmup 2 --no-cpu-baseline &&
run timeout -s KILL 90 rocprofv3 --pmc FETCH_SIZE -d "$OUT/fetch" -o plain -- \
    python3 "$REPO/bench.py" --steps 4 --warmup 0 --no-cpu-baseline &&
run timeout -s KILL 90 rocprofv3 --pmc WRITE_SIZE -d "$OUT/write" -o plain -- \
    python3 "$REPO/bench.py" --steps 4 --warmup 0 --no-cpu-baseline &&
run timeout -s KILL 60 rocprofv3 --pmc FETCH_SIZE -d "$OUT/copy_fetch" -o copy -- \
    "$REPO/ataraxia_amd/profiling/copy_probe" 268435456 &&
run timeout -s KILL 60 rocprofv3 --pmc WRITE_SIZE -d "$OUT/copy_write" -o copy -- \
    "$REPO/ataraxia_amd/profiling/copy_probe" 268435456 &&
run timeout -s KILL 90 rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_INSTS_LDS SQ_WAVE_CYCLES SQ_BUSY_CYCLES -d "$OUT/sq" -o plain -- \
    python3 "$REPO/bench.py" --steps 4 --warmup 0 --no-cpu-baseline
