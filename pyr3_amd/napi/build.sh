#!/bin/sh
# Builds pyr3hip.node (Node-API binding of libpyr3hip.so) next to this file.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
PKG=$(dirname "$HERE")
gcc -O2 -fPIC -shared -std=c11 -Wall -DNODE_GYP_MODULE_NAME=pyr3hip \
  -I/usr/include/node -I"$PKG/../include" "$HERE/pyr3_napi.c" \
  -L"$PKG" -lpyr3hip -Wl,-rpath,'$ORIGIN/..' -o "$HERE/pyr3hip.node"
