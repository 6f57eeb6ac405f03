#!/bin/bash
# Times the oracle's restatement of PARENT.x (oracle_parent.x) beside the
# reference's own PARENT.x built from its sources with the same flags, on one
# synthetic sample, so that the CPU baseline can be shown to run at the
# reference's speed.  Build container only (needs oracle/_ref from
# `make -C oracle ref`); prints "Calculation time" (PARENT.cpp:498-1489) of
# each build.
#   oracle/cpu_rates.sh [atoms] [frames] [threads]
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
ATOMS=${1:-200}
FRAMES=${2:-20000}
export OMP_NUM_THREADS=${3:-8}
WORK=$(mktemp -d)
trap 'rm -rf "$WORK"' EXIT
g++ -O3 -march=native -fopenmp -Wall -std=c++17 -o "$WORK/oracle_native.x" "$HERE/oracle_main.cpp" "$HERE/oracle.cpp"
python3 -c "import sys; sys.path.insert(0, '$ROOT'); import parent_amd as pa; pa.synth_write_bat('$WORK/s.bat', $ATOMS, $FRAMES, 12345)"
echo "sample: $ATOMS-atom chain x $FRAMES frames, 50 bins, $OMP_NUM_THREADS threads"
run() {  # label, executable, LD_LIBRARY_PATH
    local t
    t=$(LD_LIBRARY_PATH=$3 "$2" "$WORK/s.bat" "$WORK/$1.par" 50 50 50 50 50 50 | grep "Calculation time")
    echo "$1: $t"
}
run reference_shipped_flags "$HERE/_ref/PARENT.x" /opt/conda/lib
run oracle_shipped_flags "$HERE/_build/oracle_parent_shipped.x" ""
run reference_O3_native "$HERE/_ref/PARENT_native.x" /opt/conda/lib
run oracle_O3_native "$WORK/oracle_native.x" ""
run reference_O3_x86-64-v3 "$HERE/_ref/PARENT_O3.x" /opt/conda/lib
run oracle_O3_x86-64-v3 "$HERE/_build/oracle_parent_O3.x" ""
cmp "$WORK/reference_shipped_flags.par" "$WORK/oracle_shipped_flags.par" && echo "shipped-flag .par files identical"
