"""Build libpawsim_gpu.so (HIP, gfx950) and the pawsim_gpu host driver in-tree.

    python -m pawsim_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "libpawsim_gpu.so")
BIN_DIR = os.path.join(PKG, "bin")
DRIVER = os.path.join(BIN_DIR, "pawsim_gpu")

HIP_SOURCES = ["pg_kernels.hip", "pg_tendency.hip", "pg_mg.hip", "pg_mgfast.hip", "pg_net.hip", "pg_learn.hip", "pg_ensemble.hip", "pg_diag.hip", "pg_checkpoint.hip", "pg_restoring.hip", "pg_api.hip"]

# -ffp-contract=off: FP64 results must equal the serial reference bit for
# bit, so no multiply-add fusion; packed FP32 instructions are disabled.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library() -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run(["hipcc", *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", *srcs, "-o", LIB])
    return LIB


def build_driver() -> str:
    os.makedirs(BIN_DIR, exist_ok=True)
    _run(["gcc", "-O2", "-std=gnu99", "-Wall", f"-I{INCLUDE}", os.path.join(CSRC, "driver", "pawsim_gpu_main.c"),
          "-o", DRIVER, f"-L{PKG}", "-lpawsim_gpu", "-Wl,-rpath,$ORIGIN/..", "-lm"])
    return DRIVER


def build_all() -> None:
    build_library()
    build_driver()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
