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
        # the compiler's per-kernel resource remarks (registers, LDS, private
        # segment, occupancy) are kept beside the object for check_resources()
        cmd = [HIPCC, "-x", "hip", "-c", path, "-o", obj] + HIP_FLAGS + ["-Rpass-analysis=kernel-resource-usage"]
        with open(os.path.join(OBJ, src + ".resources.txt"), "w") as f:
            f.write(_run(cmd))
    else:
        cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-c", path, "-o", obj]
        _run(cmd)
    return obj


# Private-segment (scratch) budget per kernel, bytes per lane. ROCr backs a
# queue's private segment for every wavefront the device can hold and draws
# all queues' reservations from one agent-wide limit, so large per-lane
# private data on several queues at once ends in HSA_STATUS_ERROR_OUT_OF_RESOURCES.
# Lanes-per-joint kernels keep their state in registers and LDS (a few
# per-lane arrays of chain joints at most); thread-per-environment kernels hold
# one environment's per-joint arrays (up to 64 joints, FP64) per lane.
LANES_KERNELS = ("k_chain_", "k_tree_", "k_clik_lanes", "k_hetero", "k_link_state", "k_random_qpos")  # name prefixes
LANES_PRIVATE_BUDGET = 4096
PRIVATE_BUDGET = 65536
# ... and per kernel, per-lane bytes x 64 lanes x (its occupancy x 1024 SIMDs)
RESERVATION_BUDGET = 16 << 30


def resource_report():
    """Per kernel of the last build: name, source, VGPRs, AGPRs, private
    segment bytes per lane, LDS bytes per block, occupancy (waves per SIMD)."""
    import re

    rows = []
    for src in HIP_SOURCES:
        path = os.path.join(OBJ, src + ".resources.txt")
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (build first)")
        cur = None
        for line in open(path):
            m = re.search(r"Function Name: (\S+)", line)
            if m:
                cur = {"name": m.group(1), "source": src}
                rows.append(cur)
                continue
            for key, pat in (("vgpr", r"VGPRs: (\d+)"), ("agpr", r"AGPRs: (\d+)"),
                             ("private", r"ScratchSize \[bytes/lane\]: (\d+)"),
                             ("lds", r"LDS Size \[bytes/block\]: (\d+)"),
                             ("occupancy", r"Occupancy \[waves/SIMD\]: (\d+)")):
                m = re.search(pat, line)
                if m and cur is not None:
                    cur[key] = int(m.group(1))
    return rows


def kernel_base_name(mangled):
    """The unqualified function name of a mangled kernel name
    (``_ZN6sapamd12_GLOBAL__N_113k_chain_fusedILi2...`` -> ``k_chain_fused``)."""
    import re

    if not mangled.startswith("_ZN"):
        return mangled
    j, last = 3, mangled
    while j < len(mangled) and mangled[j].isdigit():
        digits = re.match(r"\d+", mangled[j:]).group()
        j += len(digits)
        last = mangled[j:j + int(digits)]
        j += int(digits)
    return last


def check_resources(rows=None):
    """Kernels over the private-segment budgets (an empty list when none is)."""
    bad = []
    for r in rows if rows is not None else resource_report():
        p = r.get("private", 0)
        base = kernel_base_name(r["name"])
        lanes = base == "k_hetero" or any(base.startswith(k) for k in LANES_KERNELS if k != "k_hetero")
        limit = LANES_PRIVATE_BUDGET if lanes else PRIVATE_BUDGET
        reserve = p * 64 * r.get("occupancy", 8) * 1024
        if p > limit or reserve > RESERVATION_BUDGET:
            bad.append((r["name"], r["source"], p, reserve))
    return bad


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
