"""Compile libnbh.so for gfx950 in place (hipcc; no JIT cache).

Flags: -ffp-contract=off and no fast-math keep the FMA error-free products and
the magic-constant round exact; packed FP32 ops are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess

from ._lib import LIB_PATH, PKG_DIR, ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["nbh_kernels.hip", "nbh_api.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def build(verbose: bool = False) -> str:
    csrc = os.path.join(PKG_DIR, "csrc")
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", csrc,
           *[os.path.join(csrc, s) for s in SOURCES], "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH
