#!/bin/bash
# Regenerates philox_rocrand.json from rocRAND's Philox4x32-10 engine.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
TMP="$(mktemp -d)"
hipcc -O1 -o "${TMP}/gen" "${HERE}/philox_rocrand_vectors.cpp"
"${TMP}/gen" > "${HERE}/philox_rocrand.json"
rm -rf "${TMP}"
