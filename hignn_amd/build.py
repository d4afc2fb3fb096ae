"""Builds libhh.so (HIP kernels + C ABI, gfx950) and the pybind11 module
``hignn`` in place, next to this file.

    python -m hignn_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIBHH = os.path.join(HERE, "libhh.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HH_ARCH", "gfx950")

# -ffp-contract=off: the tree build must round every product on its own
# (bit-exact tree vs. the oracle); fused multiply-adds in kernels are written
# explicitly.  Packed FP32 ops are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


SOURCES = ("hh_kernels.hip", "hh_forcing.hip", "hh_three_body.hip", "hh_surrogate.hip",
           "hh_tree_gpu.hip", "hh_model.cpp", "hh_tree.cpp", "hh_dynamics.cpp", "hh_three_body.cpp")


def build_libhh(jobs=None):
    """Compiles every source to an object (in parallel), then links libhh.so."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, f + ".o") for f in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, min(os.cpu_count() or 1, 16)))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, so[0]), "-o", so[1]]),
                    zip(SOURCES, objs)))
    tmp = LIBHH + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, LIBHH)
    return LIBHH


def build_pymodule():
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "hignn" + ext)
    tmp = out + ".tmp"
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"], "/opt/rocm/include"]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           *[f"-I{i}" for i in inc], os.path.join(CSRC, "hignn_py.cpp"),
           "-L" + HERE, "-lhh", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_oracle():
    """The CPU oracle is test infrastructure; it is built here so that the
    tests and the CPU-baseline leg of bench.py find it."""
    _run(["sh", os.path.join(ROOT, "oracle", "build.sh")])


def build_all():
    build_libhh()
    build_pymodule()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIBHH)
    sys.exit(0)
