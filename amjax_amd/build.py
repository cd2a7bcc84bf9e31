"""Build recipe for the native libraries (run on the host, no GPU needed).

* ``libamjax_amd.so``   - the gfx950 HIP kernels behind the C-ABI declared in
  ``include/amjax_amd.h``; compiled with ``hipcc --offload-arch=gfx950``.
* ``libamjax_setup.so`` - host-only C loops of the hierarchy setup (gcc).

Both land next to this file so they travel with the source tree.

The HIP library is compiled as several objects in parallel: the main unit and
``batch_launch.hip`` once per (value type, kernel-id unit).  ``cached=True``
(``python -m amjax_amd.build --cached``) keeps the objects in ``_build/``
keyed by a hash of the sources and flags and recompiles only those whose
inputs changed; the default always compiles everything.
"""

import hashlib
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
CACHE_DIR = os.path.join(ROOT, "_build")

HIP_LIB = os.path.join(HERE, "libamjax_amd.so")
SETUP_LIB = os.path.join(HERE, "libamjax_setup.so")

KB_NPART = 4  # must match KB_NPART in csrc/batch_launch.hpp

# (source, extra defines, object name)
HIP_UNITS = [("amjax_amd.hip", [], "amjax_amd")] + [
    ("batch_launch.hip", ["-DAMG_KB_T=" + t, "-DAMG_KB_PART=%d" % p], "batch_launch_%s_%d" % (t, p))
    for t in ("float", "double") for p in range(KB_NPART)
]
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


def _unit_hash(src, defines):
    """Hash of a unit's inputs: its source, every header it may include, flags."""
    h = hashlib.sha256()
    h.update(" ".join(HIP_FLAGS + defines).encode())
    paths = [os.path.join(CSRC, src)]
    paths += sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h")))
    paths += sorted(os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h"))
    for p in paths:
        h.update(p.encode())
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:20]


def build_hip_lib(cached=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    jobs = jobs or max(1, min(len(HIP_UNITS), os.cpu_count() or 1))
    with tempfile.TemporaryDirectory(prefix="amjax_amd_build_") as tmp:
        objs, todo, hashes = [], [], {}
        for src, defines, name in HIP_UNITS:
            if cached:
                os.makedirs(CACHE_DIR, exist_ok=True)
                hashes[name] = _unit_hash(src, defines)
                obj = os.path.join(CACHE_DIR, "%s-%s.o" % (name, hashes[name]))
                objs.append(obj)
                if os.path.exists(obj):
                    continue
            else:
                obj = os.path.join(tmp, name + ".o")
                objs.append(obj)
            cmd = [hipcc, *HIP_FLAGS, *defines, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src),
                   "-o", obj + ".tmp"]
            todo.append((cmd, obj))
        # the main unit is the longest: start it first
        running, failed = [], []
        while todo or running:
            while todo and len(running) < jobs:
                cmd, obj = todo.pop(0)
                print("+", " ".join(cmd), flush=True)
                running.append((subprocess.Popen(cmd), obj))
            finished = [(p, obj) for p, obj in running if p.poll() is not None]
            if not finished:
                time.sleep(0.2)
                continue
            for p, obj in finished:
                running.remove((p, obj))
                if p.returncode != 0:
                    failed.append(p.args)
                else:
                    os.replace(obj + ".tmp", obj)
        if failed:
            raise subprocess.CalledProcessError(1, failed[0])
        if cached:
            # a source edited while the units compiled would leave objects of mixed
            # inputs under the old keys: drop them and refuse to link
            stale = [name for src, defines, name in HIP_UNITS if _unit_hash(src, defines) != hashes[name]]
            if stale:
                for src, defines, name in HIP_UNITS:
                    if name in stale:
                        p = os.path.join(CACHE_DIR, "%s-%s.o" % (name, hashes[name]))
                        if os.path.exists(p):
                            os.remove(p)
                raise RuntimeError("sources changed during the build (%s): build again" % ", ".join(stale))
        _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", HIP_LIB + ".tmp"])
    os.replace(HIP_LIB + ".tmp", HIP_LIB)


def build_all(cached=False):
    build_setup_lib()
    build_hip_lib(cached=cached)


if __name__ == "__main__":
    args = sys.argv[1:]
    cached = "--cached" in args
    which = [a for a in args if not a.startswith("--")] or ["all"]
    for w in which:
        {"all": lambda: build_all(cached), "setup": build_setup_lib,
         "hip": lambda: build_hip_lib(cached=cached)}[w]()
