"""Builds libflgpu.so (gfx950) in-tree with hipcc, one object per source.

Run as ``python -m gnu_femlisp_amd.build`` or through __graft_entry__.build().
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libflgpu.so")
SOURCES = ["flgpu_core.hip", "flgpu_sparse.hip", "flgpu_lattice.hip", "flgpu_mg.hip",
           "flgpu_unet.hip", "flgpu_io.hip", "flgpu_block.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         # packed FP32 instructions are kept out of all device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, r.stdout + r.stderr, obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        for src, rc, out, obj in ex.map(_compile, SOURCES):
            lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if rc != 0:
                raise RuntimeError(f"hipcc failed for {src}:\n" + "\n".join(lines))
            if verbose and lines:
                print("\n".join(lines))
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
