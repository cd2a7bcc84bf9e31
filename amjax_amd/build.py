"""Build recipe for the native libraries (run on the host, no GPU needed).

* ``libamjax_amd.so``   - the gfx950 HIP kernels behind the C-ABI declared in
  ``include/amjax_amd.h``; compiled with ``hipcc --offload-arch=gfx950``.
* ``libamjax_setup.so`` - host-only C loops of the hierarchy setup (gcc).

Both land next to this file so they travel with the source tree.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIP_LIB = os.path.join(HERE, "libamjax_amd.so")
SETUP_LIB = os.path.join(HERE, "libamjax_setup.so")

HIP_SOURCES = ["amjax_amd.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    # keep the compiler from emitting packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_setup_lib():
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall",
          os.path.join(CSRC, "setup_host.c"), "-o", SETUP_LIB + ".tmp", "-lm"])
    os.replace(SETUP_LIB + ".tmp", SETUP_LIB)


def build_hip_lib():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([hipcc, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, *srcs, "-o", HIP_LIB + ".tmp"])
    os.replace(HIP_LIB + ".tmp", HIP_LIB)


def build_all():
    build_setup_lib()
    build_hip_lib()


if __name__ == "__main__":
    which = sys.argv[1:] or ["all"]
    for w in which:
        {"all": build_all, "setup": build_setup_lib, "hip": build_hip_lib}[w]()
