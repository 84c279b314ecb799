"""Build libecsprt_hip.so (gfx950) and the C host driver from source.

hipcc compiles every csrc/*.hip for --offload-arch=gfx950 and links one shared
library in-tree; gcc builds the C driver against it.  Run as
``python -m ecsprt_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libecsprt_hip.so")
DRIVER = os.path.join(HERE, "ecsprt_pic")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]
SOURCES = ["api.hip", "particles.hip", "particles_mixed.hip", "particles2d.hip", "collide.hip", "solver.hip", "batch.hip",
           "refinit.hip", "noalloc.hip", "bsolve.hip", "fileio.hip"]
# The FP64 path is built without contraction so that it reproduces the reference's
# rounding; the mixed-precision kernels are not bit-compatible by design and may fuse.
EXTRA_FLAGS = {"particles_mixed.hip": ["-ffp-contract=fast"]}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    # Keep warnings visible, minus the known host-pass note about the GPU feature flag.
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err:
        sys.stderr.write(err + "\n")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    drv_src = os.path.join(HERE, "host", "pic_main.c")
    if os.path.exists(drv_src):
        _run(["gcc", "-O2", "-std=c99", "-D_POSIX_C_SOURCE=199309L", "-Wall", f"-I{os.path.join(ROOT, 'include')}", drv_src,
              "-o", DRIVER, f"-L{HERE}", "-lecsprt_hip", "-Wl,-rpath,$ORIGIN", "-lm"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
