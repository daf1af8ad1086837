"""Autotuner for the lanes-per-joint chain kernel shape.

The fused FK + Jacobian + RNEA kernel for serial chains maps one environment to
G lanes with C joints per lane (G C >= nj). Small G spends fewer lane-cycles
per environment (less scan overhead); large G spreads a batch over more
wavefronts (more SIMDs busy, more latency hiding). The best shape depends on
the joint count, the batch size and the GPU, so it is measured: every feasible
shape is timed with HIP events on the launch stream over back-to-back launches
and the fastest is installed in the model through the C ABI
(sapamd_set_chain_shape). Results are cached per (GPU, joint count, batch) in a
JSON file so later processes skip the sweep.
"""
from __future__ import annotations

import ctypes
import json
import os

CACHE = os.environ.get("SAPAMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "sapien_amd", "tune.json"))


def shapes():
    from .native import lib

    G = (ctypes.c_int32 * 64)()
    C = (ctypes.c_int32 * 64)()
    n = lib().sapamd_chain_shapes(G, C, 64)
    return [(G[i], C[i]) for i in range(n)]


def set_shape(model, max_batch, G, C):
    from .native import check, handle, lib

    check(lib().sapamd_set_chain_shape(handle(model), ctypes.c_int64(max_batch), G, C))


def get_shape(model, B):
    from .native import check, handle, lib

    g, c = ctypes.c_int32(), ctypes.c_int32()
    check(lib().sapamd_get_chain_shape(handle(model), ctypes.c_int64(B), ctypes.byref(g), ctypes.byref(c)))
    return g.value, c.value


def _load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def autotune(model, batch_sizes=(16384,), iters=100, use_cache=True, report=None):
    """Pick and install the fastest chain-kernel shape for each batch size.

    Returns {B: (G, C, microseconds)}. Models that are not serial chains of
    <= 16 joints (where the thread-per-environment kernels run) return {}.
    """
    import torch

    from . import kernel_variant
    from .native import launcher_fk_jac_rnea

    if kernel_variant(model) != "lanes_per_joint" or not model.is_chain or model.nv > 16:
        return {}
    nj = model.nv
    dev = torch.device("cuda", torch.cuda.current_device())
    gpu = torch.cuda.get_device_name(dev)
    cache = _load_cache() if use_cache else {}
    launch = launcher_fk_jac_rnea(model)
    stream = torch.cuda.current_stream(dev)
    L = model.nlinks
    lo, hi = model.joint_limits()
    out = {}
    for B in sorted(batch_sizes):
        key = f"{gpu}|chain{nj}|L{L}|B{B}"
        if key in cache:
            G, C, us = cache[key]
            set_shape(model, B, G, C)
            out[B] = (G, C, us)
            continue
        gen = torch.Generator(device="cpu").manual_seed(B)
        q = (torch.rand(B, nj, generator=gen, dtype=torch.float64) * torch.tensor(hi - lo) + torch.tensor(lo)).float()
        q = q.to(dev)
        qd = torch.randn(B, nj, generator=gen).to(dev)
        qdd = torch.randn(B, nj, generator=gen).to(dev)
        P = torch.empty(B, L, 7, device=dev)
        J = torch.empty(B, 6, nj, device=dev)
        T = torch.empty(B, nj, device=dev)
        timings = []
        for G, C in shapes():
            if G * C < nj or (G * C) // 2 >= nj and G * C > 8:
                continue  # does not cover the chain, or wastes more than half of the lanes
            set_shape(model, B, G, C)
            for _ in range(10):
                launch(B, q, qd, qdd, P, J, T, stream.cuda_stream)
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(iters):
                launch(B, q, qd, qdd, P, J, T, stream.cuda_stream)
            b.record(stream)
            b.synchronize()
            timings.append((a.elapsed_time(b) / iters * 1e3, G, C))
            if report is not None:
                report.append({"B": B, "G": G, "C": C, "us": timings[-1][0]})
        us, G, C = min(timings)
        set_shape(model, B, G, C)
        out[B] = (G, C, us)
        cache[key] = [G, C, us]
    if use_cache:
        _save_cache(cache)
    return out
