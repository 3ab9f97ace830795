"""Build recipe of libppm_hip.so (hipcc, gfx950), in-tree next to this file.

    python -m spammm_amd.build          # compile every source, link the library

Each source is compiled to an object under spammm_amd/_build/ and the objects are
linked into spammm_amd/libppm_hip.so.  The relaxation kernels are compiled with
-ffp-contract=off (their FP32 operation order is part of their contract), and so is
the FP64 reference forward (the operation order of the FP64 CPU oracle).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libppm_hip.so")
ARCH = os.environ.get("PPM_OFFLOAD_ARCH", "gfx950")

# (source, extra flags)
SOURCES = [
    ("ppm_api.hip", []),
    ("ppm_grid.hip", ["-ffp-contract=off"]),
    ("ppm_relax.hip", ["-ffp-contract=off"]),
    ("ppm_df.hip", ["-ffp-contract=fast"]),
    ("ppm_batch.hip", ["-ffp-contract=off"]),
    ("ppm_direct.hip", ["-ffp-contract=off"]),
    ("ppm_f64.hip", ["-ffp-contract=off"]),
    ("ppm_fdbm.hip", []),
    ("ppm_ewald.hip", []),
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (set HIPCC)")


def common_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        "-Wall", "-Wno-unused-function",
    ]


def _compile(src_flags):
    src, extra = src_flags
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + common_flags() + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lhipfft"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
