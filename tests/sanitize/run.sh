#!/bin/bash
# Host library and oracle under the sanitizers, on the CPU (no GPU needed):
#  1. UBSan builds of libucuda.so (host code) and liboracle.so in a copy of
#     the working tree, and the CPU test suite (pytest -m "not gpu") on them;
#  2. ASan + UBSan executables: a C-ABI driver (systems, codegen, step
#     bodies, request validation on a device-less engine), the session-layer
#     translation unit tests/session_fields.cpp ("cpu"), and an oracle driver
#     running every analysis on small inputs. Leak detection on.
# Python loads the sanitized libraries late, which only UBSan tolerates, so
# ASan runs on the standalone executables. Usage: tests/sanitize/run.sh [workdir]
set -euo pipefail
REPO=$(cd "$(dirname "$0")/../.." && pwd)
W=${1:-$(mktemp -d /tmp/ucuda_san.XXXXXX)}
ROCM=/opt/rocm
S=$W/tree
mkdir -p "$S"
# the working tree without history, results or working files
(cd "$REPO" && tar --exclude=./.git --exclude=./_work --exclude=./gpurun_out --exclude='*.so' \
   --exclude=./u-cuda_amd/kcache -cf - .) | tar -xf - -C "$S"
cd "$S/u-cuda_amd"
python3 -c "import build; build.write_embed(); build.build_scratch_object()"
C=$S/u-cuda_amd/csrc
SRCS="$C/engine.cpp $C/codegen.cpp $C/parametric_engine.cpp $C/data_export.cpp"
NOFUSE="-fno-builtin-sin -fno-builtin-cos -fno-builtin-sinf -fno-builtin-cosf"
SANRT=$(dirname "$($ROCM/llvm/bin/clang -print-file-name=libclang_rt.ubsan_standalone-x86_64.so)")
# 1. UBSan libraries and the CPU suite
$ROCM/bin/hipcc -O1 -g -std=c++17 -fPIC -shared $NOFUSE \
  -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=undefined -shared-libsan \
  -o "$S/u-cuda_amd/libucuda.so" $SRCS -x none "$S/u-cuda_amd/scratch.o" \
  -L$ROCM/lib -lhiprtc -lamdhip64 -ldl -Wl,-rpath,$ROCM/lib -Wl,-rpath,"$SANRT"
gcc -O1 -g -ffp-contract=off -fno-fast-math $NOFUSE -fPIC -fopenmp -std=gnu11 \
  -fsanitize=undefined -fno-sanitize-recover=undefined -shared -o "$S/oracle/liboracle.so" "$S/oracle/ucuda_oracle.c" -lm
(cd "$S" && UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 python3 -m pytest tests/ -x -q -m "not gpu" -p no:cacheprovider)
# 2. ASan + UBSan executables
SAN="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer"
for drv in "$REPO/tests/sanitize/abi_driver.cpp" "$REPO/tests/session_fields.cpp"; do
  $ROCM/bin/hipcc -O1 -g -std=c++17 $NOFUSE $SAN -I"$S/include" -I"$C" -o "$W/$(basename "$drv" .cpp)" $SRCS "$drv" \
    -x none "$S/u-cuda_amd/scratch.o" -L$ROCM/lib -lhiprtc -lamdhip64 -ldl -Wl,-rpath,$ROCM/lib
done
gcc -O1 -g -fno-omit-frame-pointer -ffp-contract=off -fopenmp -std=gnu11 -fsanitize=address,undefined \
  -fno-sanitize-recover=undefined -I"$S/oracle" -o "$W/oracle_driver" "$REPO/tests/sanitize/oracle_driver.c" \
  "$S/oracle/ucuda_oracle.c" -lm
export ASAN_OPTIONS=detect_leaks=1:halt_on_error=1 UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
"$W/abi_driver"
"$W/session_fields" cpu
"$W/oracle_driver"
echo "sanitizers: clean"
