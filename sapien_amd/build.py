"""Builds the native parts of sapien_amd in-tree for gfx950.

* ``libsapien_amd.so`` -- HIP kernels + the C ABI of ``include/sapien_amd.h``
  (hipcc, one object per source, linked with -shared);
* ``_pinocchio*.so``   -- the pybind11 host module (g++), linked to the C ABI
  library through an ``$ORIGIN`` rpath.

Every HIP translation unit is built with packed FP32 disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).

Usage: ``python sapien_amd/build.py`` (by path: importing the package needs the
extension this builds), or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SAPAMD_ARCH", "gfx950")

HIP_SOURCES = ["kernels_generic.hip", "kernels_fast.hip", "kernels_adjoint.hip", "kernels_chain_vjp.hip", "kernels_tree.hip", "kernels_rollout.hip", "kernels_rollout_tree.hip", "kernels_clik_lanes.hip", "abi.cpp"]
HOST_SOURCES = ["model.cpp"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include")]
LIB = os.path.join(HERE, "libsapien_amd.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip") or src in ("abi.cpp",):
        cmd = [HIPCC, "-x", "hip", "-c", path, "-o", obj] + HIP_FLAGS
    else:
        cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_lib(jobs=4):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_module():
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_pinocchio" + suffix)
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"], os.path.join(ROOT, "include")]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", os.path.join(CSRC, "pymodule.cpp"),
           "-o", out + ".tmp"] + ["-I" + i for i in inc] + ["-L" + HERE, "-lsapien_amd", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build():
    build_lib()
    build_module()


if __name__ == "__main__":
    build()
    print("built", LIB)
