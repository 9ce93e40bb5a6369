#!/bin/bash
# AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU test suite, on
# the CPU only (no GPU is touched):
#   1. the oracle built by gcc with both sanitizers, under the oracle, pin and
#      golden tests;
#   2. libsvg_amd.so with the sanitizers on its host code only (each
#      -fsanitize= behind -Xarch_host; the device code is built as usual),
#      under the host-side mesh, geodesic and ABI tests.
# Outputs go to tests/_san/ (not committed).  Usage: tests/sanitize_host.sh
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT=$ROOT/tests/_san
HIPCC=/opt/rocm/bin/hipcc
mkdir -p "$OUT/obj"

echo "== oracle (gcc, ASan + UBSan)"
gcc -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=undefined \
    -fPIC -std=gnu11 -ffp-contract=off -fopenmp -shared -o "$OUT/libsvg_oracle.so" \
    "$ROOT/oracle/svg_oracle.c" -lm
GCC_RT="$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so)"
(cd "$ROOT" && SVG_ORACLE_LIB="$OUT/libsvg_oracle.so" LD_PRELOAD="$GCC_RT" \
    ASAN_OPTIONS=detect_leaks=0:abort_on_error=1 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
    python -m pytest tests/test_oracle.py tests/test_oracle_pins.py tests/test_oracle_point.py \
        tests/test_golden.py -q -x -m "not gpu" -p no:cacheprovider)

echo "== libsvg_amd.so host code (clang, ASan + UBSan)"
SAN="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer"
COMMON="-O1 -g -std=c++17 -fPIC -ffp-contract=off -I $ROOT/include -I $ROOT/svg_amd/csrc"
for f in svg_mesh.cpp geodesic.cpp capi.cpp; do
    $HIPCC $COMMON $SAN -c "$ROOT/svg_amd/csrc/$f" -o "$OUT/obj/$f.o"
done
for f in build.hip solve.hip serve.hip point.hip; do
    $HIPCC -x hip --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops -munsafe-fp-atomics \
        $COMMON $SAN -c "$ROOT/svg_amd/csrc/$f" -o "$OUT/obj/$f.o"
done
$HIPCC -shared -fPIC --offload-arch=gfx950 $SAN -Xarch_host -shared-libsan -o "$OUT/libsvg_amd.so" \
    "$OUT"/obj/*.o -lpthread
CLANG_RT=$(ls /opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -n 1)
(cd "$ROOT" && SVG_AMD_LIB="$OUT/libsvg_amd.so" LD_PRELOAD="$CLANG_RT" \
    ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:protect_shadow_gap=0 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
    python -m pytest tests/test_mesh_host.py tests/test_geodesic_host.py tests/test_abi.py \
        -q -x -m "not gpu" -p no:cacheprovider)
echo "sanitized runs clean"
