#!/bin/sh
# oracle/dropin.sh -- TEST INFRASTRUCTURE: the reference's own benchmark with
# its optimisable files replaced by hpcg_amd/host/, built the way a
# maintainer would build the drop-in.
#
#   dropin.sh build REF_SRC OUT    the reference's main.cpp, TestCG,
#                                  TestSymmetry, ReportResults, ... linked
#                                  against libhpcg_amd.so -> OUT
#   dropin.sh leak  REF_SRC        the reference's generator
#                                  (GenerateProblem_ref, GenerateCoarseProblem,
#                                  SetupHalo) with the replacement headers,
#                                  built with AddressSanitizer: generates a
#                                  4-level 16^3 problem, frees it with
#                                  DeleteMatrix / DeleteCGData / DeleteVector
#                                  and fails on any leak
#
# The reference sources are copied into a temporary directory outside the
# repository (a quoted #include finds the including file's own directory
# first, so the replacement headers must sit beside the sources), the
# replacement files are copied over them, and only the binary is written
# to OUT.  Every hpcg_amd/host file replaces or joins the reference's
# except Geometry.hpp: the reference's own is kept (the driver's copy holds
# only what the driver uses).
set -eu
mode=$1
REFSRC=$2
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
ROCM=${ROCM:-/opt/rocm}
CXX=${CXX:-g++}
JOBS=${JOBS:-8}
tmp=$(mktemp -d)
trap 'rm -rf "$tmp"' EXIT

cp "$REFSRC"/*.cpp "$REFSRC"/*.hpp "$tmp"/
for f in "$ROOT"/hpcg_amd/host/*.cpp "$ROOT"/hpcg_amd/host/*.hpp; do
  case $(basename "$f") in Geometry.hpp) continue ;; esac
  cp "$f" "$tmp"/
done

FLAGS="-O3 -fopenmp -DHPCG_NO_MPI -std=c++17 -w -D__HIP_PLATFORM_AMD__ -I$ROCM/include -I$ROOT/include"
LIBS="-L$ROOT/hpcg_amd -lhpcg_amd -L$ROCM/lib -lamdhip64"

case $mode in
build)
  OUT=$3
  (cd "$tmp" && ls *.cpp | xargs -P "$JOBS" -I{} sh -c "$CXX $FLAGS -c {} -o {}.o")
  mkdir -p "$(dirname "$OUT")"
  # rpath: the in-tree library, found relative to OUT (oracle/_ref/ -> hpcg_amd/)
  $CXX -fopenmp -o "$OUT.tmp" "$tmp"/*.o $LIBS \
    -Wl,-rpath,'$ORIGIN/../../hpcg_amd' -Wl,-rpath,"$ROCM/lib"
  mv "$OUT.tmp" "$OUT"
  ;;
leak)
  rm "$tmp/main.cpp"
  cp "$HERE/leak_check.cpp" "$tmp"/
  ASAN="-O1 -g -fsanitize=address -fno-omit-frame-pointer"
  (cd "$tmp" && ls *.cpp | xargs -P "$JOBS" -I{} sh -c "$CXX $ASAN $FLAGS -c {} -o {}.o")
  $CXX -fsanitize=address -fopenmp -o "$tmp/leak_check" "$tmp"/*.o $LIBS \
    -Wl,-rpath,"$ROOT/hpcg_amd" -Wl,-rpath,"$ROCM/lib"
  ASAN_OPTIONS=detect_leaks=1 "$tmp/leak_check"
  ;;
*)
  echo "usage: $0 build REF_SRC OUT | leak REF_SRC" >&2
  exit 2
  ;;
esac
