#!/bin/bash
# Build liborf_rcwa.so (gfx950) in-tree.  Used by __graft_entry__.build().
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${HERE}/liborf_rcwa.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(--offload-arch=gfx950 -O3 -fPIC -std=c++17 -Xclang -target-feature -Xclang -packed-fp32-ops
       -munsafe-fp-atomics -Wno-unused-result)
mkdir -p "${HERE}/_build"
objs=()
pids=()
for src in la_kernels rcwa_kernels orf_plan; do
  rm -f "${HERE}/_build/${src}.o"
  "${HIPCC}" "${FLAGS[@]}" -c "${HERE}/csrc/${src}.hip" -o "${HERE}/_build/${src}.o" &
  pids+=($!)
  objs+=("${HERE}/_build/${src}.o")
done
for p in "${pids[@]}"; do wait "$p"; done
"${HIPCC}" --offload-arch=gfx950 -shared -o "${OUT}.tmp" "${objs[@]}"
mv "${OUT}.tmp" "${OUT}"
echo "built ${OUT}"
