"""Build the qvmhip HIP library in-tree: qvm_amd/_lib/libqvmhip.so (gfx950).

Every source is compiled by hipcc for --offload-arch=gfx950 and linked into
one shared library whose only public symbols are the extern "C" entry points
of include/qvmhip.h.  Packed-FP32 instructions are disabled for all device
code.  The plain gate, tile and reduction kernels are compiled with
-ffp-contract=off so each complex product is rounded like the reference's
serial kernels; the fused MFMA kernel keeps contraction.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
LIB = os.path.join(OUT_DIR, "libqvmhip.so")

ARCH = "gfx950"
SOURCES = {
    "api.hip": [],
    "gates.hip": ["-ffp-contract=off"],
    "tile.hip": ["-ffp-contract=off"],
    "reduce.hip": ["-ffp-contract=off"],
    "fused.hip": [],
    "program.hip": [],
    "mlp.hip": [],
    "batch.hip": [],
    "noalloc.hip": [],
}
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
    "-I", os.path.join(REPO, "include"),
]


STAMP = os.path.join(OUT_DIR, "libqvmhip.stamp")


def source_hash() -> str:
    """SHA-256 of everything the library is built from: the HIP sources,
    the headers and the compile flags."""
    import hashlib
    h = hashlib.sha256()
    files = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".h")))
    files += [os.path.join(REPO, "include", f) for f in sorted(os.listdir(os.path.join(REPO, "include")))]
    for path in files:
        h.update(os.path.relpath(path, REPO).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    flags = [f.replace(REPO, "<repo>") for f in COMMON]  # the tree may sit anywhere
    h.update(repr((ARCH, flags, sorted(SOURCES.items()))).encode())
    return h.hexdigest()


def hipcc() -> str:
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if p and os.path.exists(p):
            return p
    return "hipcc"


def _compile(src: str, extra: list) -> str:
    obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
    cmd = [hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    with open(STAMP, "w") as f:
        f.write(source_hash() + "\n")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
