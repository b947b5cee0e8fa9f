#!/bin/sh
# Copies the input.yml of each sim directory under the reference's
# testing/netflux_tests/ into sim_inputs/<dir>.yml (configuration data only;
# the reference ships no meshes for them).
set -e
src=${1:-/root/reference/testing/netflux_tests}
dst=$(dirname "$0")/sim_inputs
mkdir -p "$dst"
for d in "$src"/*/; do
    n=$(basename "$d")
    [ -f "$d/input.yml" ] && cp "$d/input.yml" "$dst/$n.yml"
done
