"""Builds libtaichirtm_hip.so (HIP, gfx950) in-tree next to this file.

    python -m taichirtm_amd.build

Flags: -ffp-contract=off keeps the reference's float32 expression order
operand by operand (no FMA contraction), so kernel results are bitwise
reproducible; packed FP32 instructions are disabled for the device code.
"""
import concurrent.futures
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtaichirtm_hip.so")
SOURCES = ["trtm_abi.hip", "trtm_simple.hip", "trtm_fused.hip", "trtm_block.hip",
           "trtm_adjoint.hip", "trtm_fwi_block.hip", "trtm_wave.hip"] + [
    "trtm_wave_%s_%s.hip" % (sys_, dir_) for sys_ in ("sh", "psv", "both")
    for dir_ in ("fwd", "bwd")]
ARCH = os.environ.get("TRTM_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def compile_command(src, obj):
    return [hipcc(), *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c",
            os.path.join(CSRC, src), "-o", obj]


def link_command(objs, out):
    return [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, "-o", out]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: %s" % " ".join(cmd[-4:]))


def build(verbose=False):
    """Compiles every source to an object (in parallel), then links the
    library; the previous library is replaced only when all steps succeed."""
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory(prefix="trtm_build_") as work:
        objs = [os.path.join(work, os.path.splitext(src)[0] + ".o") for src in SOURCES]
        jobs = min(len(SOURCES), max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
            for f in [pool.submit(_run, compile_command(src, obj), verbose)
                      for src, obj in zip(SOURCES, objs)]:
                f.result()
        _run(link_command(objs, tmp), verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
