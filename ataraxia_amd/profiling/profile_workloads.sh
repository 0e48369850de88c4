#!/bin/bash
# rocprofv3 kernel-trace passes behind profiles/serve_poisson200_kernels.txt
# and profiles/batch_mix_kernels.txt.  Run from the repo root on a machine
# with the GPU:
#     [OUT=dir] bash ataraxia_amd/profiling/profile_workloads.sh
# then summarise each with: python ataraxia_amd/profiling/summarize.py <OUT>/<workload>
#   serve:     the serving loop, 256 Poisson requests at 200 requests/s
#   batch_mix: the 64-scene batch through the batched renderer only
# Kernel trace and stats only (no counters); every pass has its own time
# limit and the script stops at the first failure.
set -u
REPO=$(pwd)
OUT=${OUT:-$REPO/profile_out/workloads}
case "$OUT" in /*) ;; *) OUT="$REPO/$OUT" ;; esac   # the passes run from /tmp
mkdir -p "$OUT"
export TMPDIR=/tmp
run() { echo "== $*"; "$@"; local rc=$?; echo "== rc=$rc"; return $rc; }
cd /tmp || exit 1
run timeout -k 10 240 rocprofv3 --kernel-trace --stats -d "$OUT/serve/kt" -o k -- \
    python3 "$REPO/bench.py" --workload serve --rate 200 --requests 256 &&
run timeout -k 10 240 rocprofv3 --kernel-trace --stats -d "$OUT/batch_mix/kt" -o k -- \
    python3 "$REPO/bench.py" --workload batch_mix --no-plain
