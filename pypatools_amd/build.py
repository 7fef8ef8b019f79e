"""Build libpypat_hip.so in-tree for gfx950 with hipcc.

    python -m pypatools_amd.build            # or __graft_entry__.build()

Each translation unit is compiled on its own (the strict TU must not see FMA
contraction), then linked into pypatools_amd/libpypat_hip.so.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(PKG, "libpypat_hip.so")
BUILD_DIR = os.path.join(PKG, "_build")

ARCH = os.environ.get("PYPAT_OFFLOAD_ARCH", "gfx950")
SOURCES = ["pp_host.hip", "pp_strict.hip", "pp_fast.hip", "pp_sort.hip", "pp_moments.hip", "pp_collide.hip",
           "pp_poisson.hip", "pp_spline.hip"]
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-I", INCLUDE, "-I", CSRC,
    # packed FP32 (v_pk_*_f32) stays off: scalar FP32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# the fast-path kernel sources: their sha256 is compiled into the library (pp_kernel_digest), so that
# profiler counters taken on one build are matched to the binary that runs, not to the files on disk
KERNEL_SOURCES = ("pp_fast.hip", "pp_internal.h", "pp_fp64.inc", "pp_bvh.inc")


def kernel_source_digest():
    """sha256 over the fast-path kernel sources as they are on disk."""
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(CSRC, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


PER_FILE = {
    "pp_strict.hip": ["-ffp-contract=off"],
    "pp_collide.hip": ["-ffp-contract=off"],
    "pp_poisson.hip": ["-ffp-contract=off"],
    "pp_spline.hip": ["-ffp-contract=off"],
}


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return exe


def _compile(src, digest):
    obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
    extra = [f'-DPP_KERNEL_DIGEST="{digest}"'] if src == "pp_host.hip" else []
    cmd = [hipcc(), "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + PER_FILE.get(src, []) + extra
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-6000:]}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD_DIR, exist_ok=True)
    digest = kernel_source_digest()
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda src: _compile(src, digest), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} ({os.path.getsize(OUT) // 1024} KiB) for {ARCH}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
