"""Build recipe for ``libbioshell_gpu.so`` (gfx950 only).

Every HIP and C++ source under ``csrc/`` is compiled by ``hipcc`` for
``--offload-arch=gfx950`` and linked into one shared library next to this
file.  Packed-FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
DIAG = os.path.join(HERE, "diag")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbioshell_gpu.so")
# peak-rate probes and host builds of the device QCP code, for the bench and
# the CPU tests; kept out of the product library
DIAG_LIB = os.path.join(HERE, "libbioshell_diag.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources(where: str = CSRC) -> list[str]:
    return sorted(
        os.path.join(where, f) for f in os.listdir(where) if f.endswith((".hip", ".cpp"))
    )


def _compile(src: str) -> str:
    sub = "diag_" if os.path.dirname(src) == DIAG else ""
    obj = os.path.join(OBJ, sub + os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON_FLAGS, "-I", CSRC, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def _link(objs: list[str], lib: str) -> None:
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp, *objs, "-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, lib)


def build(jobs: int | None = None) -> str:
    """Compiles all sources and links the product library (and the
    diagnostics library); returns the product library's path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs, diag = sources(), sources(DIAG)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs + diag))
    _link(objs[:len(srcs)], LIB)
    _link(objs[len(srcs):], DIAG_LIB)
    return LIB


if __name__ == "__main__":
    print(build())
