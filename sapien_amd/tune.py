"""Autotuner for the lanes-per-joint kernel shapes.

* Chains: the fused FK + Jacobian + RNEA kernel for serial chains maps one
  environment to G lanes with C joints per lane (G C >= nj). Small G spends
  fewer lane-cycles per environment (less scan overhead); large G spreads a
  batch over more wavefronts (more SIMDs busy, more latency hiding).
  Installed with sapamd_set_chain_shape.
* Trees and the mixed-tree launch: lanes per environment (8, 16 or 32, at
  least the joint count) of every model's segment, timed over the whole
  mixed launch (sapamd_set_tree_group).
* CLIK: the library-wide schedule of the lanes-per-joint CLIK launches
  (first-phase iteration cap, lanes per problem in each phase), timed over
  the caller's CLIK workload (sapamd_set_clik_schedule).

The best shape depends on the joint counts, the batch sizes and the GPU, so it
is measured: every candidate is timed (HIP events on the launch stream, or the
host clock around a synchronized call) and the fastest is installed through
the C ABI. Results are cached per GPU and workload in a JSON file so later
processes skip the sweep.
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


def chain_candidates(nj):
    """The (lanes per environment, joints per lane) shapes the autotuner tries
    for a chain of nj joints: they cover the chain and waste less than half
    the lanes."""
    return [(G, C) for G, C in shapes() if not (G * C < nj or (G * C) // 2 >= nj and G * C > 8)]


def autotune(model, batch_sizes=(16384,), iters=100, use_cache=True, report=None):
    """Pick and install the fastest chain-kernel shape for each batch size.

    Each candidate is timed by its kernels' own durations (kernel_us: launches
    queued without host gaps, each recording its kernel's start and stop into
    HIP events; host and dispatch overheads, a profiler's included, do not
    count) over launches that rotate through enough input/output sets to miss
    the Infinity Cache, three rounds over all candidates, best round each.
    Returns {B: (G, C, microseconds)}. Models that are not serial chains of
    <= 16 joints (where the thread-per-environment kernels run) return {}.
    """
    import torch

    from . import kernel_variant
    from .native import launcher_fk_jac_rnea, launcher_fk_jac_rnea_timed

    if kernel_variant(model) != "lanes_per_joint" or not model.is_chain or model.nv > 16:
        return {}
    nj = model.nv
    dev = torch.device("cuda", torch.cuda.current_device())
    gpu = torch.cuda.get_device_name(dev)
    cache = _load_cache() if use_cache else {}
    launch = launcher_fk_jac_rnea(model)
    timed = launcher_fk_jac_rnea_timed(model)
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
        cands = chain_candidates(nj)
        best = {}
        for _ in range(3):  # rounds over all candidates; the best round of each counts
            for G, C in cands:
                set_shape(model, B, G, C)
                for i in range(10):
                    launch(B, *sets[i % nsets], stream.cuda_stream)
                t = kernel_us(lambda i, e0, e1: timed(B, *sets[i % nsets], e0, e1, stream.cuda_stream), iters)
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


# ---------------------------------------------------------------- tree kernels and mixed-tree launches
def tree_groups(model):
    """Lanes-per-environment choices of the lanes-per-joint tree kernels for this model."""
    nj = model.nv
    return [G for G in (8, 16, 32) if G >= nj]


def set_tree_group(model, G):
    from .native import check, handle, lib

    check(lib().sapamd_set_tree_group(handle(model), int(G)))


def get_tree_group(model):
    from .native import check, handle, lib

    g = ctypes.c_int32()
    check(lib().sapamd_get_tree_group(handle(model), ctypes.byref(g)))
    return g.value


def _gpu_name():
    import torch

    return torch.cuda.get_device_name(torch.device("cuda", torch.cuda.current_device()))


_SLEEP_MS_PER_MCYCLE = None


def hold_gpu(ms):
    """Keep the current stream's GPU busy for about `ms` milliseconds (a sleep
    kernel), so that the launches the host queues meanwhile wait on the GPU,
    not on the host."""
    import torch

    global _SLEEP_MS_PER_MCYCLE
    if _SLEEP_MS_PER_MCYCLE is None:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        torch.cuda._sleep(1 << 20)
        b.record()
        b.synchronize()
        _SLEEP_MS_PER_MCYCLE = max(a.elapsed_time(b), 1e-3) / ((1 << 20) / 1e6)
    torch.cuda._sleep(int(min(ms / _SLEEP_MS_PER_MCYCLE, 4e3) * 1e6))


_EVENTS = []


def kernel_us(run_timed, iters):
    """Mean kernel duration in microseconds of `iters` launches queued behind a
    sleep kernel (so the host's launch rate does not space them out);
    run_timed(i, start, stop) makes launch i recording its kernel's own start
    and stop into the two raw event handles. Host launch overhead and the
    dispatch gaps between launches are not counted (nor a profiler's)."""
    import torch

    while len(_EVENTS) < iters:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()  # torch creates an event at its first record
        b.record()
        _EVENTS.append((a, b))
    torch.cuda.synchronize()
    hold_gpu(2.0 + 0.05 * iters)
    for i in range(iters):
        run_timed(i, _EVENTS[i][0].cuda_event, _EVENTS[i][1].cuda_event)
    torch.cuda.synchronize()
    return 1e3 * sum(_EVENTS[i][0].elapsed_time(_EVENTS[i][1]) for i in range(iters)) / iters


def _time(fn, iters, rounds=3):
    """Best of `rounds` HIP-event timings of `iters` back-to-back calls, in microseconds."""
    import torch

    stream = torch.cuda.current_stream()
    for _ in range(3):
        fn()
    best = float("inf")
    for _ in range(rounds):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(iters):
            fn()
        b.record(stream)
        b.synchronize()
        best = min(best, a.elapsed_time(b) / iters * 1e3)
    return best


def autotune_hetero(models, batches, iters=50, use_cache=True, report=None):
    """Pick the lanes per environment of every model's segment of the mixed-tree
    FK + J + RNEA + CRBA launch (sapamd_hetero_batched) by timing every
    combination of the models' choices over the given per-model batch sizes
    (q, qd, qdd uniform in [-1, 1]); installs the fastest through
    sapamd_set_tree_group and returns ({model index: G}, microseconds)."""
    import itertools

    import torch

    from .native import launcher_hetero

    dev = torch.device("cuda", torch.cuda.current_device())
    key = f"{_gpu_name()}|hetero|" + "|".join(f"{m.nv}x{m.nlinks}x{b}" for m, b in zip(models, batches))
    cache = _load_cache() if use_cache else {}
    if key in cache:
        gs, us = cache[key]
        for m, G in zip(models, gs):
            set_tree_group(m, G)
        return dict(enumerate(gs)), us
    gen = torch.Generator(device="cpu").manual_seed(7)
    ins = [[(torch.rand(b, m.nv, generator=gen) * 2 - 1).to(dev) for _ in range(3)] for m, b in zip(models, batches)]
    outs = [(torch.empty(b, m.nlinks, 7, device=dev), torch.empty(b, 6, m.nv, device=dev),
             torch.empty(b, m.nv, device=dev), torch.empty(b, m.nv, m.nv, device=dev)) for m, b in zip(models, batches)]
    stream = torch.cuda.current_stream(dev)
    best = None
    for gs in itertools.product(*[tree_groups(m) for m in models]):
        for m, G in zip(models, gs):
            set_tree_group(m, G)
        launch = launcher_hetero(models, batches, [x[0] for x in ins], [x[1] for x in ins], [x[2] for x in ins],
                                 *[[o[k] for o in outs] for k in range(4)])
        us = _time(lambda: launch(stream.cuda_stream), iters)
        if report is not None:
            report.append({"lanes_per_env": list(gs), "us": us})
        if best is None or us < best[1]:
            best = (list(gs), us)
    for m, G in zip(models, best[0]):
        set_tree_group(m, G)
    cache[key] = best
    if use_cache:
        _save_cache(cache)
    return dict(enumerate(best[0])), best[1]


# ---------------------------------------------------------------- CLIK schedule
CLIK_SCHEDULES = [(128, 4, 8), (64, 4, 8), (256, 4, 8), (128, 4, 16), (128, 8, 16), (128, 4, 4), (0, 4, 8), (0, 8, 8)]


def set_clik_schedule(cap, g1, g2):
    from .native import check, lib

    check(lib().sapamd_set_clik_schedule(int(cap), int(g1), int(g2)))


def get_clik_schedule():
    from .native import check, lib

    c, a, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    check(lib().sapamd_get_clik_schedule(ctypes.byref(c), ctypes.byref(a), ctypes.byref(b)))
    return c.value, a.value, b.value


def autotune_clik(run, tag, candidates=CLIK_SCHEDULES, rounds=2, use_cache=True, report=None):
    """Pick the library-wide lanes-per-joint CLIK schedule (first-phase
    iteration cap, lanes per problem in each phase) by timing `run()` (one
    batched CLIK call over the caller's workload, e.g. the mixed-tree batch)
    under every candidate; installs the fastest and returns ((cap, g1, g2),
    milliseconds). A problem's result bits depend on the schedule (the
    association order of its sums), not on the batch it runs in."""
    import time

    import torch

    key = f"{_gpu_name()}|clik|{tag}"
    cache = _load_cache() if use_cache else {}
    if key in cache:
        sch, ms = cache[key]
        set_clik_schedule(*sch)
        return tuple(sch), ms
    best = None
    for sch in candidates:
        set_clik_schedule(*sch)
        run()
        torch.cuda.synchronize()
        t = float("inf")
        for _ in range(rounds):
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            t = min(t, (time.perf_counter() - t0) * 1e3)
        if report is not None:
            report.append({"cap": sch[0], "g1": sch[1], "g2": sch[2], "ms": t})
        if best is None or t < best[1]:
            best = (sch, t)
    set_clik_schedule(*best[0])
    cache[key] = [list(best[0]), best[1]]
    if use_cache:
        _save_cache(cache)
    return tuple(best[0]), best[1]
