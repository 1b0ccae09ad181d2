"""Build libzkgpu_ntt.so in-tree with hipcc for gfx950.

    python -m zk_gpu_amd.build [--no-force]

Objects go to zk_gpu_amd/_build/, the library to zk_gpu_amd/libzkgpu_ntt.so.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libzkgpu_ntt.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ZKGPU_ARCH", "gfx950")

SOURCES = ["ntt_kernels.hip", "noalloc.hip", "msm_kernels.hip", "batch.cpp", "engine.cpp", "capi.cpp", "serve.cpp",
           "host_stage.cpp", "msm.cpp"]
# packed-FP32 instructions are disabled for all device code (they are not
# used by this integer code, and some of their forms are unreliable)
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src, force):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + ".o")
    if not force and os.path.exists(out):
        deps = [path] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith((".hpp", ".h"))]
        deps.append(os.path.join(HERE, "..", "include", "zkgpu_ntt.h"))
        if all(os.path.getmtime(out) >= os.path.getmtime(d) for d in deps if os.path.exists(d)):
            return out
    cmd = [HIPCC] + FLAGS + ["-I", CSRC, "-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(force=True, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    # -z defs: an undefined symbol fails the link instead of the first load
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,-z,defs", "-o", tmp] + objs + ["-lpthread", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-force", action="store_true")
    a = ap.parse_args()
    build(force=not a.no_force, verbose=True)
    sys.exit(0)
