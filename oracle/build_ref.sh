#!/usr/bin/env bash
# Build the reference CPU program (pba.out) from its single translation unit
# under /root/reference/src, into oracle/_ref/ (git-ignored).  Only used to
# write the golden fixtures and as the "reference" CPU baseline.
# HDF5 1.10 (C++ API) comes from the image's /opt/conda; only the HDF5 shared
# objects and their compression/transport dependencies are exposed through a
# private symlink folder, so that conda's older libstdc++ is not picked up.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC=/root/reference/src/main.cpp
OUT="$HERE/_ref"
CONDA=/opt/conda
if [ ! -f "$SRC" ]; then
  echo "reference source not present; skipping reference build" >&2
  exit 0
fi
if [ ! -f "$CONDA/include/H5Cpp.h" ]; then
  echo "HDF5 C++ headers not present in $CONDA; reference is unbuildable here" >&2
  exit 0
fi
mkdir -p "$OUT/lib"
for pat in 'libhdf5*.so*' 'libsz.so*' 'libaec.so*' 'libz.so*' 'libcurl.so*' 'libssl.so*' 'libcrypto.so*'; do
  for f in $CONDA/lib/$pat; do
    [ -e "$f" ] && ln -sf "$f" "$OUT/lib/$(basename "$f")"
  done
done
g++ "$SRC" -o "$OUT/pba.out" -O2 -std=c++17 -fopenmp -w \
    -I"$CONDA/include" -L"$OUT/lib" -Wl,-rpath,"$OUT/lib" \
    -lhdf5_cpp -lhdf5 -lstdc++fs -lpthread
echo "built $OUT/pba.out"
