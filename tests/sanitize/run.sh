#!/bin/bash
# Host-side AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU test
# suite (no GPU needed; device code is not instrumented):
#   1. the oracle and the CPU baseline built by gcc with -fsanitize=address,undefined,
#      under gcc's runtimes;
#   2. the host code of libb64hip (C ABI, domain setup, FRI parameters, transcript,
#      NTT planner, proof serialization) built by hipcc with the sanitizers on the
#      host pass only, under clang's runtime.
# Builds go to tests/sanitize/_build; the test loaders pick them up through
# B64_SANITIZED_LIBS.  Usage: bash tests/sanitize/run.sh
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
OUT=$ROOT/tests/sanitize/_build
mkdir -p "$OUT/gcc" "$OUT/hip"
cd "$ROOT"

GSAN="-O1 -g -fPIC -std=gnu11 -mpclmul -msse4.1 -fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer"
gcc $GSAN -shared -o "$OUT/gcc/libb64oracle.so" oracle/b64_oracle.c -lpthread -lm
gcc $GSAN -shared -o "$OUT/gcc/libb64cpu.so" oracle/b64_cpu_baseline.c -lpthread -lm
ln -sf "$ROOT/binius64_amd/libb64hip.so" "$OUT/gcc/libb64hip.so"
echo "== oracle and CPU baseline under ASan + UBSan"
LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)" \
ASAN_OPTIONS=detect_leaks=0 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
B64_SANITIZED_LIBS="$OUT/gcc" python -m pytest tests -x -q -m "not gpu" -p no:cacheprovider

HSAN="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=undefined -Xarch_host -fno-omit-frame-pointer"
pids=()
for src in api kernels batch server pcs vision; do
    /opt/rocm/bin/hipcc -O1 -g -std=c++17 -fPIC --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
        -I include $HSAN -c "binius64_amd/csrc/$src.hip" -o "$OUT/hip/$src.o" 2> "$OUT/hip/$src.log" & pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
/opt/rocm/bin/hipcc -shared -fPIC --offload-arch=gfx950 -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined \
    -Xarch_host -shared-libsan "$OUT"/hip/*.o -o "$OUT/hip/libb64hip.so"
ln -sf "$ROOT/oracle/libb64oracle.so" "$OUT/hip/libb64oracle.so"
ln -sf "$ROOT/oracle/libb64cpu.so" "$OUT/hip/libb64cpu.so"
RT=$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)
echo "== libb64hip host code under ASan + UBSan"
LD_PRELOAD="$RT" ASAN_OPTIONS=detect_leaks=0 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
B64_SANITIZED_LIBS="$OUT/hip" python -m pytest tests -x -q -m "not gpu" -p no:cacheprovider
