"""Build recipe for the native libraries (run on the host, no GPU needed).

* ``libamjax_amd.so``   - the gfx950 HIP kernels behind the C-ABI declared in
  ``include/amjax_amd.h``; compiled with ``hipcc --offload-arch=gfx950``.
* ``libamjax_setup.so`` - host-only C loops of the hierarchy setup (gcc).

Both land next to this file so they travel with the source tree.
"""

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIP_LIB = os.path.join(HERE, "libamjax_amd.so")
SETUP_LIB = os.path.join(HERE, "libamjax_setup.so")

# compiled in parallel, one object each, then linked into one library
HIP_SOURCES = ["amjax_amd.hip", "batch_launch.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
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
    with tempfile.TemporaryDirectory(prefix="amjax_amd_build_") as tmp:
        objs, procs = [], []
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            cmd = [hipcc, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
            print("+", " ".join(cmd), flush=True)
            procs.append(subprocess.Popen(cmd))
            objs.append(obj)
        failed = [p.args for p in procs if p.wait() != 0]
        if failed:
            raise subprocess.CalledProcessError(1, failed[0])
        _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", HIP_LIB + ".tmp"])
    os.replace(HIP_LIB + ".tmp", HIP_LIB)


def build_all():
    build_setup_lib()
    build_hip_lib()


if __name__ == "__main__":
    which = sys.argv[1:] or ["all"]
    for w in which:
        {"all": build_all, "setup": build_setup_lib, "hip": build_hip_lib}[w]()
