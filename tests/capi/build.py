"""Builds tests/capi/capi_check (a C++ caller of the library's C ABI) with g++ against the
built libmarshal_amd/libmarshal_hip.so. The binary finds the library through a relative
rpath, so it runs from any copy of the tree."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "capi_check")


def build(out=OUT):
    cmd = ["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "capi_check.cpp"),
           "-L" + os.path.join(ROOT, "libmarshal_amd"), "-lmarshal_hip", "-L/opt/rocm/lib",
           "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../libmarshal_amd", "-Wl,-rpath,/opt/rocm/lib",
           "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"capi_check build failed:\n{r.stdout}\n{r.stderr}")
    return out


if __name__ == "__main__":
    print(build())
