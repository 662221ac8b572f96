"""Build libnbabfs.so (HIP, gfx950) in-tree with hipcc.

``python -m pdynamo_amd.build`` or ``build()`` compiles every source under
``pdynamo-mirror_amd/csrc`` in parallel and links ``pdynamo_amd/libnbabfs.so``.
"""

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "libnbabfs.so")
SOURCES = ["nbabfs_build.hip", "nbabfs_force.hip", "nbabfs_host.hip", "nbabfs_qcmm.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")
# Packed FP32 instructions stay disabled; the kernels use scalar FP32 and FP64.
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + INCLUDE, "-I" + CSRC]


def _compile(src, out):
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-4000:]))
    return out


def build(verbose=False):
    """Compile and link libnbabfs.so; returns its path."""
    tmp = tempfile.mkdtemp(prefix="nbabfs_build_")
    try:
        objs = [os.path.join(tmp, s.replace(".hip", ".o")) for s in SOURCES]
        with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
            list(ex.map(_compile, SOURCES, objs))
        out_tmp = LIB + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", out_tmp]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
        os.replace(out_tmp, LIB)
        if verbose:
            print("built", LIB)
        return LIB
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
