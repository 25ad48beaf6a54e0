#!/usr/bin/env bash
# Builds oracle/_ref/libsdft_ref.so from the reference's own sources where they
# lie under /root/reference (read-only), plus oracle/ref_driver.cpp. The
# reference's vendored Eigen 3.4 keeps its module headers loose and its src/
# tree zipped; both are staged into a temporary directory under oracle/_ref and
# removed after compiling. Nothing from the reference is committed; the output
# .so is git-ignored. Test infrastructure only (pins of the CPU oracle).
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
STAGE="$OUT/_eigen_stage"
[ -d "$REF/src" ] || { echo "reference sources not present; skipping"; exit 0; }
mkdir -p "$OUT"
rm -rf "$STAGE"
mkdir -p "$STAGE/Eigen"
cp -r "$REF/third_party/eigen3/Eigen/." "$STAGE/Eigen/"
rm -f "$STAGE/Eigen/src.zip"
python3 -c "import zipfile,sys; zipfile.ZipFile(sys.argv[1]).extractall(sys.argv[2])" \
  "$REF/third_party/eigen3/Eigen/src.zip" "$STAGE/Eigen"
SRCS=(pw_core.cpp compute_backend.cpp band_group_solver.cpp rmm_trial.cpp indefinite_pencil_davidson.cpp sdft_core.cpp)
FLAGS=(-std=c++20 -O2 -fPIC -fopenmp -ffp-contract=off -w -DSDFT_HAS_EIGEN=1 -DEIGEN_MPL2_ONLY -DSDFT_HAS_OPENMP=1
       -I"$STAGE" -I"$REF/src")
OBJS=()
pids=()
for s in "${SRCS[@]}"; do
  o="$STAGE/${s%.cpp}.o"
  g++ "${FLAGS[@]}" -c "$REF/src/$s" -o "$o" & pids+=($!)
  OBJS+=("$o")
done
g++ "${FLAGS[@]}" -c "$HERE/ref_driver.cpp" -o "$STAGE/ref_driver.o" & pids+=($!)
OBJS+=("$STAGE/ref_driver.o")
for p in "${pids[@]}"; do wait "$p"; done
g++ -shared -fopenmp "${OBJS[@]}" -o "$OUT/libsdft_ref.so.tmp" -lcrypto
mv "$OUT/libsdft_ref.so.tmp" "$OUT/libsdft_ref.so"
rm -rf "$STAGE"
echo "built $OUT/libsdft_ref.so"
