#!/bin/sh
# Regenerates the golden fixtures in this directory by running the oracle
# harness (oracle/ref_harness/harness.cpp compiled against the reference's own
# headers).  Needs the reference tree (REFERENCE, default /root/reference).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
make -C "$ROOT/oracle" ref REFERENCE="${REFERENCE:-/root/reference}"
"$ROOT/oracle/_ref/harness" "$HERE"
