"""Builds the in-tree HIP library esperanto_amd/lib/libesp_align.so for gfx950.

    python -m esperanto_amd.build          # or esperanto_amd.build.build()

hipcc cross-compiles here without a GPU; the built .so travels with the tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libesp_align.so")
SOURCES = ["esp_kernels.hip", "esp_band.hip", "esp_full.hip", "esp_myers.hip", "esp_window.hip",
           "esp_walk.hip", "esp_wave.hip", "esp_host.cpp", "esp_window.cpp", "esp_walk.cpp", "esp_server.cpp",
           "esp_graph.cpp", "esp_noalloc.hip", "esp_junction.hip", "esp_junction.cpp",
           "esp_synth.cpp"]
ARCH = os.environ.get("ESP_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # Packed FP32 instructions stay off in every HIP object of this library.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-pthread",
]


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libesp_align.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
