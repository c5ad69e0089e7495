"""Builds libmarshal_amd/libmarshal_hip.so for gfx950 with hipcc (no CMake, no JIT cache).

Every HIP source under csrc/ is compiled to an object in csrc/_build/ and linked into one
shared library next to this file, so the library travels with the source tree.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmarshal_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["kernels_010.hip", "kernels_100.hip", "kernels_rc.hip", "kernels_pad.hip", "kernels_st.hip", "kernels_seg.hip", "kernels_cl.hip", "batch.hip",
           "serve.hip", "api.hip", "noalloc.hip", "ktiming.hip"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # Packed-FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _compile(src):
    obj = os.path.join(CSRC, "_build", os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    # compiler warnings are shown (the host pass's note that it does not know the packed-FP32
    # feature is expected and left out)
    warnings = [ln for ln in r.stderr.splitlines()
                if "warning:" in ln and "packed-fp32-ops" not in ln]
    return obj, warnings


def build(verbose=False):
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        done = list(ex.map(_compile, srcs))
    objs = [o for o, _ in done]
    for _, warnings in done:
        for w in warnings:
            print(w, file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
