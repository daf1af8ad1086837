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

    Each candidate is timed with HIP events over launches that rotate through
    enough input/output sets to miss the Infinity Cache, three rounds over all
    candidates, best round each. Returns {B: (G, C, microseconds)}. Models that
    are not serial chains of <= 16 joints (where the thread-per-environment
    kernels run) return {}.
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
        key = f"{gpu}|chain{nj}|L{L}|B{B}|rotated"
        if key in cache:
            G, C, us = cache[key]
            set_shape(model, B, G, C)
            out[B] = (G, C, us)
            continue
        # input/output sets rotated so the working set (>= 512 MB) exceeds the
        # 256 MB Infinity Cache, as in a stream of fresh batches
        per_env = 4 * (3 * nj + 7 * L + 6 * nj + nj)
        nsets = max(2, min(64, -(-(512 << 20) // (per_env * B))))
        gen = torch.Generator(device="cpu").manual_seed(B)
        sets = []
        for _ in range(nsets):
            q = (torch.rand(B, nj, generator=gen, dtype=torch.float64) * torch.tensor(hi - lo) + torch.tensor(lo))
            sets.append((q.float().to(dev), torch.randn(B, nj, generator=gen).to(dev),
                         torch.randn(B, nj, generator=gen).to(dev), torch.empty(B, L, 7, device=dev),
                         torch.empty(B, 6, nj, device=dev), torch.empty(B, nj, device=dev)))
        cands = [(G, C) for G, C in shapes()
                 if not (G * C < nj or (G * C) // 2 >= nj and G * C > 8)]  # covers the chain, wastes < half the lanes
        best = {}
        for _ in range(3):  # rounds over all candidates; the best round of each counts
            for G, C in cands:
                set_shape(model, B, G, C)
                for i in range(10):
                    launch(B, *sets[i % nsets], stream.cuda_stream)
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record(stream)
                for i in range(iters):
                    launch(B, *sets[i % nsets], stream.cuda_stream)
                b.record(stream)
                b.synchronize()
                t = a.elapsed_time(b) / iters * 1e3
                best[(G, C)] = min(best.get((G, C), t), t)
        del sets
        timings = [(t, G, C) for (G, C), t in best.items()]
        if report is not None:
            report.extend({"B": B, "G": G, "C": C, "us": t} for t, G, C in timings)
        us, G, C = min(timings)
        set_shape(model, B, G, C)
        out[B] = (G, C, us)
        cache[key] = [G, C, us]
    if use_cache:
        _save_cache(cache)
    return out
