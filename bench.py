"""Benchmark: batched FK (all link poses) + full Jacobian + RNEA evals/s at
B = 16384 environments of the seeded 7-dof chain, FP32, one MI355X.

One step = one launch of the fused kernel over one batch of 16384 environments,
inputs already resident in HBM. Step i uses input/output buffer set i % 64
(64 sets, ~530 MB, more than the 256 MB Infinity Cache). After `warmup`
launches and a 512 MiB read that evicts the Infinity Cache, the timed region
is `steps` calls of the C-ABI entry point sapamd_fk_jac_rnea_batched (ctypes,
device pointers resolved beforehand), bracketed by HIP events on the launch
stream and by the host clock. `value` = 16384 x steps / host time. Set-up
launches the same kernel before the warm-up: the shape autotuner, one pass
over every buffer set (the reference outputs the timed launches are checked
against) and the launch-mode diagnostics (the first timed sequence after
set-up ran 20-60 % slower than later ones, whatever the launch mode). Prints
ONE JSON line.

Beside the headline number it reports:
  roofline      algorithmic bytes per launch / mean launch duration (HIP events
                on the launch stream around the timed region, per launch),
                against 8 TB/s HBM3E; traffic from rocprofv3
                FETCH_SIZE + WRITE_SIZE;
                the FP32 VALU work (rocprofv3 SQ_INSTS_VALU_* counters) against
                the 157.3 TFLOP/s vector peak; the bandwidth an elementwise
                copy reaches on this GPU;
  cpu_baseline  the FP64 C++ oracle (same algorithms) timed on this host's cores
                on a bounded sample, on one core, and through its per-call
                Python API (one configuration per call);
  launch_modes  the same launches captured in a HIP graph, each kernel's own
                start and stop, and the tensor-level Python launcher;
  plain_gpu     the drop-in per-call path (one configuration per call, FP64 on
                the GPU, host copies) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FK+Jacobian+RNEA evals/s @16384 envs 7-DoF; fwd+bwd grad evals/s vs plain"
B = 16384
NSETS = 64
BYTES_PER_EVAL = 4 * (3 * 7 + 7 * 8 + 6 * 7 + 7)  # q, qd, qdd in; 8 poses x 7, J 6 x 7, tau 7 out = 504
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", default="auto", help="auto | generic | fast (kernel variant of the fused path)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes")
    ap.add_argument("--no-workloads", action="store_true", help="skip the mixed-tree CLIK / rollout / hetero sections")
    ap.add_argument("--workloads-only", action="store_true",
                    help="run only the mixed-tree CLIK / rollout / hetero sections (for profiling them)")
    ap.add_argument("--shape", type=int, nargs=2, default=None, metavar=("G", "C"),
                    help="lanes-per-joint shape instead of autotuning (profiling runs, where tracing skews the tuner)")
    ap.add_argument("--probe-traffic", action="store_true", help=argparse.SUPPRESS)
    return ap.parse_args()


COPY_BYTES = 256 << 20  # calibration copy: 256 MiB read + 256 MiB written
# FP32 VALU work of the fused kernel, one rocprofv3 --pmc pass (wavefront
# instruction counts, and the FLOP count the SQ derives per instruction)
VALU_COUNTERS = ("SQ_INSTS_VALU_FLOPS_FP32", "SQ_INSTS_VALU_FLOPS_FP32_TRANS", "SQ_INSTS_VALU_FMA_F32",
                 "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_VALU")
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32, vendor figure


def probe_traffic(kernel="auto"):
    """Child of a rocprofv3 --pmc pass: 16 launches of the fused kernel over 16
    buffer sets for every lanes-per-joint shape the autotuner may pick (the
    parent keeps the figures of the one it picks), then 5 elementwise passes
    over 256 MiB (read 256 MiB, write 256 MiB) for the counter calibration."""
    import numpy as np
    import torch

    import sapien_amd as sa
    from sapien_amd import synthetic as S
    from sapien_amd import tune

    dev = torch.device("cuda:0")
    rob = S.chain7()
    m = sa.PinocchioModel(rob.urdf, [0.0, 0.0, -9.81])
    m.set_joint_order(rob.joint_names)
    m.set_link_order(rob.link_names)
    if kernel != "auto":
        sa.set_kernel_variant(m, kernel)
    shapes = tune.chain_candidates(rob.nv) if sa.kernel_variant(m) == "lanes_per_joint" else [None]
    # 16 rotating buffer sets (132 MB, beyond the 32 MB of L2) so that each
    # launch's inputs come from outside L2, as in the timed run
    sets = []
    for k in range(16):
        q, qd, qdd, _ = (torch.from_numpy(x).to(dev) for x in S.samples(rob, B, 1001 + k, dtype=np.float32))
        sets.append((q, qd, qdd, torch.empty(B, rob.nlinks, 7, device=dev), torch.empty(B, 6, rob.nv, device=dev),
                     torch.empty(B, rob.nv, device=dev)))
    launch = sa.native_fk_jac_rnea(m)
    stream = torch.cuda.current_stream(dev)
    for shape in shapes:
        if shape is not None:
            tune.set_shape(m, B, shape[0], shape[1])
        for k in range(16):
            launch(B, *sets[k], stream.cuda_stream)
    src = torch.ones(COPY_BYTES // 4, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.neg(src, out=dst)  # an elementwise kernel: reads 256 MiB, writes 256 MiB
    torch.cuda.synchronize()


def measure_traffic(kernel="auto", timeout=180):
    """HBM bytes per launch of the fused kernel, per kernel instantiation (every
    shape the autotuner may pick), from two rocprofv3 --pmc passes
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by the factor
    that makes the same counter read 256 MiB for an elementwise kernel that
    reads and writes 256 MiB, and a pass of SQ VALU counters. Each pass is a
    child process (rocprofv3 -> python3 bench.py --probe-traffic), started
    before this process touches the GPU. Returns {kernel name: figures}, or
    None if rocprofv3 is missing or fails."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile

    if shutil.which("rocprofv3") is None:
        return None
    res = {}
    for counters in (("FETCH_SIZE",), ("WRITE_SIZE",), VALU_COUNTERS):
        counter = counters[0]
        d = tempfile.mkdtemp(prefix="sapamd_pmc_")
        cmd = ["rocprofv3", "--pmc", *counters, "-d", d, "-o", "pmc", "--output-format", "csv", "--",
               sys.executable, os.path.abspath(__file__), "--probe-traffic", "--kernel", kernel]
        p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, start_new_session=True)
        try:
            p.wait(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            return None
        kern, copy = {}, []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    cname = row.get("Counter_Name")
                    if cname not in counters:
                        continue
                    name, v = row.get("Kernel_Name", ""), float(row.get("Counter_Value", "nan"))
                    if "k_chain_fused" in name or "k_kinematics" in name or "k_tree_fused" in name:
                        kern.setdefault(name, {}).setdefault(cname, []).append(v)
                    elif "neg" in name:
                        copy.append(v)
        shutil.rmtree(d, ignore_errors=True)
        med = lambda v: sorted(v)[len(v) // 2]
        if counters is VALU_COUNTERS:
            for name, c in kern.items():
                if all(c.get(x) for x in VALU_COUNTERS):
                    res.setdefault(name, {})["VALU"] = {x: med(c[x]) for x in VALU_COUNTERS}
            continue
        if not kern or not copy:
            return None
        scale = (COPY_BYTES / 1024.0) / med(copy)
        for name, c in kern.items():
            kib = med(c[counter])
            res.setdefault(name, {})[counter] = {"kib_raw": kib, "copy_kib_raw": med(copy), "scale": scale,
                                                 "bytes": kib * 1024.0 * scale}
    return res


def traffic_of(traffic, shape, variant):
    """The figures of the kernel the timed run used, from measure_traffic."""
    if not traffic:
        return None
    key = f"k_chain_fused<{shape[0]}, {shape[1]}>" if shape else ("k_kinematics" if variant != "lanes_per_joint"
                                                                  else "k_tree_fused")
    for name, v in traffic.items():
        if key in name and "FETCH_SIZE" in v and "WRITE_SIZE" in v:
            return v
    return None


def mixed_workloads(sa, S, np, torch, dev):
    """The mixed batch of the four trees: FK+J+RNEA+CRBA of 8192 environments
    in one launch, 65,536 CLIK problems in one persistent launch, and 4096 ABA
    rollouts x 128 steps (dt = 1/240, checkpoint every 16) forward and backward."""
    import time as _t

    from sapien_amd import tune

    MIX, _split = S.MIX, S.split_mix
    models, robs = [], []
    for name, _ in MIX:
        rob = S.robot(name)
        m = sa.PinocchioModel(rob.urdf, [0.0, 0.0, -9.81])
        m.set_joint_order(rob.joint_names)
        m.set_link_order(rob.link_names)
        models.append(m)
        robs.append(rob)
    res = {}

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = _t.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (_t.perf_counter() - t0) / reps

    # heterogeneous FK + J + RNEA + CRBA, one launch
    nb = _split(8192)
    qs, qds, qdds = [], [], []
    for i, (rob, b) in enumerate(zip(robs, nb)):
        q, qd, qdd, _ = S.samples(rob, b, 1002 + i, dtype=np.float32)
        qs.append(torch.from_numpy(q).to(dev))
        qds.append(torch.from_numpy(qd).to(dev))
        qdds.append(torch.from_numpy(qdd).to(dev))
    # lanes per environment of every tree's segment, autotuned over this launch
    hrep = []
    hshape, _ = tune.autotune_hetero(models, nb, use_cache=False, report=hrep)
    sec = timed(lambda: sa.hetero_batched(models, qs, qds, qdds), 50)
    # the same launch through the C ABI over preallocated outputs (poses, J,
    # tau, full symmetric M), HIP events on the launch stream
    from sapien_amd.native import launcher_hetero

    outs_h = [(torch.empty(b, rob.nlinks, 7, device=dev), torch.empty(b, 6, rob.nv, device=dev),
               torch.empty(b, rob.nv, device=dev), torch.empty(b, rob.nv, rob.nv, device=dev))
              for rob, b in zip(robs, nb)]
    hl = launcher_hetero(models, nb, qs, qds, qdds, *[[o[k] for o in outs_h] for k in range(4)])
    hs = torch.cuda.current_stream(dev)
    for _ in range(20):
        hl(hs.cuda_stream)
    h0, h1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = _t.perf_counter()
    h0.record(hs)
    for _ in range(500):
        hl(hs.cuda_stream)
    h1.record(hs)
    torch.cuda.synchronize()
    wall = (_t.perf_counter() - t0) / 500
    kms = h0.elapsed_time(h1) / 500
    # algorithmic bytes: q, qd, qdd in; poses, J, tau and the upper triangle of M out
    hbytes = sum(b * 4 * (3 * r.nv + 7 * r.nlinks + 6 * r.nv + r.nv + r.nv * (r.nv + 1) // 2)
                 for r, b in zip(robs, nb))
    res["hetero_fk_jac_rnea_crba"] = {
        "envs": nb, "evals_s": 8192 / wall, "ms": wall * 1e3, "kernel_ms": kms,
        "bytes_per_launch": hbytes, "achieved_gbs": hbytes / (kms * 1e-3) / 1e9,
        "frac": hbytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "python_api_ms": sec * 1e3,
        "shapes": {name: {"lanes_per_env": hshape[i], "envs_per_wavefront": 64 // hshape[i]}
                   for i, (name, _) in enumerate(MIX)},
        "autotuned": True, "tuning": hrep,
        "note": "one launch through the C ABI over preallocated outputs (M written in full, the byte count takes "
                "its upper triangle); python_api_ms = the Python method, outputs allocated per call"}

    # CLIK: 65,536 problems, targets = FK of random configurations, q0 = neutral
    nc = _split(65536)
    tg = []
    for i, (m, rob, b) in enumerate(zip(models, robs, nc)):
        q = torch.from_numpy(S.samples(rob, b, 1004 + i, dtype=np.float32)[0]).to(dev)
        tg.append(m.compute_forward_kinematics_batched(q)[:, rob.ee_link].contiguous())
    links = [rob.ee_link for rob in robs]
    crep = []
    sched, _ = tune.autotune_clik(lambda: sa.clik_hetero_batched(models, links, tg), "mixed65536",
                                  use_cache=False, report=crep)
    outs = sa.clik_hetero_batched(models, links, tg)
    torch.cuda.synchronize()
    t0 = _t.perf_counter()
    outs = sa.clik_hetero_batched(models, links, tg)
    torch.cuda.synchronize()
    sec = _t.perf_counter() - t0
    succ = float(sum(o[1].sum().item() for o in outs)) / 65536
    iters = float(sum(o[3].sum().item() for o in outs)) / 65536
    res["clik_mixed"] = {"problems": nc, "problems_s": 65536 / sec, "ms": sec * 1e3, "success_rate": succ,
                         "mean_iterations": iters, "max_iterations": 1000,
                         "schedule": {"phase1_iterations": sched[0], "phase1_lanes_per_problem": sched[1],
                                      "phase1_problems_per_wavefront": 64 // sched[1],
                                      "phase2_lanes_per_problem": sched[2] if sched[0] else None,
                                      "autotuned": True},
                         "tuning": crep,
                         "kernel": "lanes per joint, persistent launches with per-problem early exit: the first runs "
                                   "the phase-1 iterations, the second the problems still running"}

    # rollouts: 4096 split 4:3:2:1, 128 steps, tau ~ N(0, var 5), forward and backward
    nr = _split(4096)
    T, K, dt = S.ROLLOUT_STEPS, S.ROLLOUT_CKPT, S.ROLLOUT_DT
    per = []
    for i, (m, rob, b) in enumerate(zip(models, robs, nr)):
        q0, qd0, tau = S.rollout_workload(rob, b, 1005 + i)
        per.append((m, torch.from_numpy(q0).to(dev), torch.from_numpy(qd0).to(dev), torch.from_numpy(tau).to(dev)))
    ck = {}
    # one default-priority stream per tree: the trees' step launches overlap
    streams = [torch.cuda.Stream(device=dev) for _ in per]
    main = torch.cuda.current_stream(dev)

    def fwd_one(j):
        m, q0, qd0, tau = per[j]
        ck[j] = m.rollout_batched(q0, qd0, tau, dt, K)

    def bwd_one(j):
        m, q0, qd0, tau = per[j]
        g = torch.ones_like(q0)
        m.rollout_backward_batched(tau, ck[j][2], dt, g, g, K)

    def seq(fn):
        return lambda: [fn(j) for j in range(len(per))]

    def conc(fn):
        def run():
            for j, st in enumerate(streams):
                st.wait_stream(main)
                with torch.cuda.stream(st):
                    fn(j)
            for st in streams:
                main.wait_stream(st)
        return run

    # all four trees in one launch forward and one backward
    mq0, mqd0, mtau = [p[1] for p in per], [p[2] for p in per], [p[3] for p in per]
    mix = {}

    def fwd_mixed():
        mix["f"] = sa.rollout_hetero_batched(models, mq0, mqd0, mtau, dt, K)

    fwd_mixed()
    mones = [torch.ones_like(x) for x in mq0]
    mck = [f[2] for f in mix["f"]]

    def bwd_mixed():
        sa.rollout_backward_hetero_batched(models, mtau, mck, dt, mones, mones, K)

    mf, mb = timed(fwd_mixed, 5), timed(bwd_mixed, 5)
    sf, sb = timed(seq(fwd_one), 3), timed(seq(bwd_one), 3)
    cf, cb = timed(conc(fwd_one), 3), timed(conc(bwd_one), 3)
    trees = {}
    for j, (name, _) in enumerate(MIX):
        tf, tb = timed(lambda: fwd_one(j), 3), timed(lambda: bwd_one(j), 3)
        trees[name] = {"rollouts": nr[j], "forward_ms": tf * 1e3, "backward_ms": tb * 1e3,
                       "precision": sa.dynamics_precision(models[j])[0],
                       "kernel": "lanes per joint" if sa.kernel_variant(models[j]) == "lanes_per_joint"
                       and robs[j].nv <= 32 else "thread per env",
                       # share of rollouts whose final state is finite (explicit Euler at this
                       # step is unstable for quad18's light feet, in FP64 as well)
                       "final_state_finite": float(torch.isfinite(ck[j][0]).all(dim=1).float().mean())}
    res["rollouts_mixed"] = {"rollouts": nr, "steps": T, "dt": dt, "checkpoint_every": K,
                             "forward_env_steps_s": 4096 * T / mf, "forward_ms": mf * 1e3,
                             "backward_env_steps_s": 4096 * T / mb, "backward_ms": mb * 1e3,
                             "kernel": "all four trees in one launch forward and one backward (T steps per "
                                       "launch, state in registers, segments recomputed inside the adjoint)",
                             "one_launch_per_tree_on_four_streams": {"forward_ms": cf * 1e3, "backward_ms": cb * 1e3},
                             "one_launch_per_tree_sequential": {"forward_ms": sf * 1e3, "backward_ms": sb * 1e3},
                             "per_tree": trees}
    return res


def main():
    args = parse()
    if args.probe_traffic:
        probe_traffic(args.kernel)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: sapien_amd runs on a single MI355X "
              "(the environments shard trivially across GPUs; no multi-GPU run is made).", file=sys.stderr)
        sys.exit(2)
    # the counter passes (child processes) before this process touches the GPU
    all_traffic = None if (args.no_traffic or args.workloads_only) else measure_traffic(args.kernel)
    import numpy as np
    import torch

    import sapien_amd as sa
    from sapien_amd import synthetic as S

    dev = torch.device("cuda:0")
    if args.workloads_only:
        print(json.dumps({"workloads": mixed_workloads(sa, S, np, torch, dev)}))
        return
    rob = S.chain7()
    m = sa.PinocchioModel(rob.urdf, [0.0, 0.0, -9.81])
    m.set_joint_order(rob.joint_names)
    m.set_link_order(rob.link_names)
    if args.kernel != "auto":
        sa.set_kernel_variant(m, args.kernel)
    n, L = rob.nv, rob.nlinks
    from sapien_amd import tune

    if args.shape and sa.kernel_variant(m) == "lanes_per_joint":
        tune.set_shape(m, B, args.shape[0], args.shape[1])
        tuned = {B: (args.shape[0], args.shape[1], None)}
    else:
        tuned = tune.autotune(m, batch_sizes=(B,), use_cache=False) if sa.kernel_variant(m) == "lanes_per_joint" else {}

    # 64 rotating buffer sets, seeded like the workload (seed 1001 + set index)
    sets = []
    for k in range(NSETS):
        q, qd, qdd, _ = S.samples(rob, B, 1001 + k, dtype=np.float32)
        sets.append(dict(q=torch.from_numpy(q).to(dev), qd=torch.from_numpy(qd).to(dev),
                         qdd=torch.from_numpy(qdd).to(dev),
                         P=torch.zeros(B, L, 7, device=dev), J=torch.zeros(B, 6, n, device=dev),
                         T=torch.zeros(B, n, device=dev)))  # outputs written once at set-up
    stream = torch.cuda.current_stream(dev)
    launch = sa.native_fk_jac_rnea(m)  # C-ABI launcher over tensors (set-up and diagnostics)

    def step(i, st=None):
        s = sets[i % NSETS]
        launch(B, s["q"], s["qd"], s["qdd"], s["P"], s["J"], s["T"], st or stream.cuda_stream)

    # The timed region's launches: sapamd_fk_jac_rnea_batched called through
    # ctypes with the device pointers resolved beforehand, as a C or cgo caller
    # holds them; one call = one launch over one buffer set.
    from sapien_amd import native

    abi = native.lib().sapamd_fk_jac_rnea_batched
    handle = native.handle(m)
    calls = [(handle, B) + tuple(sets[i % NSETS][k].data_ptr() for k in ("q", "qd", "qdd", "P", "J", "T")) +
             (stream.cuda_stream,) for i in range(args.steps)]

    def run_calls(cs):
        rcs = [abi(*c) for c in cs]
        for rc in rcs:
            native.check(rc)

    # 512 MiB read between the warm-up and the timed region: the 256 MB
    # Infinity Cache then holds none of the timed launches' buffers (and no
    # dirty lines of its own to write back)
    flush = torch.ones(128 << 20, dtype=torch.float32, device=dev)

    # Set-up pass: one launch over every buffer set. Its outputs are the
    # reference that the timed launches are checked against bit for bit (the
    # timed sets' outputs are NaN-filled before the timed region).
    for i in range(NSETS):
        step(i)
    torch.cuda.synchronize()
    outs = [(sets[i]["P"], sets[i]["J"], sets[i]["T"]) for i in range(min(args.steps, NSETS))]
    refs = [[t.clone() for t in o] for o in outs]
    flat_outs = [t for o in outs for t in o]

    def nan_fill():
        # one multi-tensor launch (x NaN) rather than a fill launch per tensor:
        # 60 small launches from Python just before the timed region made it
        # 5-30 % slower on the host clock, one launch did not
        torch._foreach_mul_(flat_outs, float("nan"))

    def same_as_setup():
        return all(torch.equal(a, b) for o, r in zip(outs, refs) for a, b in zip(r, o))

    # the timing events exist before the timed region (torch creates an
    # event at its first record)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    e1.record(stream)

    # Not the metric: the same `steps` launches captured once in a HIP graph
    # and replayed after the same cache flush. Its first replay carries the
    # graph's one-time launch cost, and kernels inside a replay run slower
    # than the same launches on the stream; reported, first replay and the
    # median of three later ones.
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream(dev)
    cap.wait_stream(stream)
    with torch.cuda.graph(graph, stream=cap):
        cs = torch.cuda.current_stream(dev).cuda_stream
        for i in range(args.steps):
            step(i, cs)
    stream.wait_stream(cap)
    torch.cuda.synchronize()

    def timed_seq(fn):
        flush.sum()
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - s0) / args.steps, e0.elapsed_time(e1) / args.steps

    nan_fill()
    graph_first = timed_seq(graph.replay)
    graph_ok = same_as_setup()
    if not graph_ok:
        raise RuntimeError("the graph replay differs from the set-up pass")
    later = sorted(timed_seq(graph.replay) for _ in range(3))
    graph_later = later[1]

    # each kernel's own start and stop (hipExtLaunchKernel events) over the
    # timed region's launches queued behind a sleep kernel; these
    # end when the kernel's writes are issued, before the end-of-kernel cache
    # release a trace counts, so they read lower (reported, not the roofline's)
    timed_launch = sa.native_fk_jac_rnea_timed(m)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in evs:  # torch creates an event at its first record
        a.record(stream)
        b.record(stream)
    flush.sum()
    tune.hold_gpu(2.0 + 0.05 * args.steps)
    for i, (a, b) in enumerate(evs):
        s = sets[i % NSETS]
        timed_launch(B, s["q"], s["qd"], s["qdd"], s["P"], s["J"], s["T"], a.cuda_event, b.cuda_event,
                     stream.cuda_stream)
    torch.cuda.synchronize()
    kernel_durations = [a.elapsed_time(b) for a, b in evs]
    ext_ms = sum(kernel_durations) / len(kernel_durations)

    # not the metric: the same launches through the tensor-level Python
    # launcher (device pointers looked up per call)
    eager_wall, eager_ms = timed_seq(lambda: [step(i) for i in range(args.steps)])

    # The launch-mode diagnostics above ran first, so the GPU and the host
    # are past their first-use costs when the timed region starts.
    # warm-up: `warmup` launches over the sets counted down from the last one
    # (with few steps, not the timed sets)
    run_calls([(handle, B) + tuple(sets[NSETS - 1 - (i % NSETS)][k].data_ptr()
                                   for k in ("q", "qd", "qdd", "P", "J", "T")) + (stream.cuda_stream,)
               for i in range(args.warmup)])
    nan_fill()
    flush.sum()
    torch.cuda.synchronize()
    # The timed region: exactly `steps` launches, step i over buffer set i % 64
    t0 = time.perf_counter()
    e0.record(stream)
    run_calls(calls)
    e1.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # the roofline's launch duration: HIP events over the timed region, per launch
    kern_ms = e0.elapsed_time(e1) / args.steps
    if not same_as_setup():
        raise RuntimeError("the timed launches differ from the set-up pass")
    value = B * args.steps / elapsed
    achieved = BYTES_PER_EVAL * B / (kern_ms * 1e-3) / 1e9

    del flush
    traffic = traffic_of(all_traffic, tuned[B][:2] if tuned else None, sa.kernel_variant(m))

    # achievable HBM bandwidth on this GPU: an elementwise pass over 256 MiB
    # (read + write), the calibration kernel of the traffic counters
    csrc = torch.ones(COPY_BYTES // 4, device=dev)
    cdst = torch.empty_like(csrc)
    for _ in range(3):
        torch.neg(csrc, out=cdst)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    for _ in range(10):
        torch.neg(csrc, out=cdst)
    c1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 2 * COPY_BYTES * 10 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del csrc, cdst

    launch_modes = {
        "c_abi": {"ms_per_step": elapsed / args.steps * 1e3, "events_ms_per_launch": kern_ms,
                  "matches_setup_pass": True,
                  "what": f"the timed region: {args.steps} calls of sapamd_fk_jac_rnea_batched through ctypes with "
                          f"device pointers resolved beforehand, one launch each"},
        "graph": {"first_replay": {"ms_per_step": graph_first[0] * 1e3, "events_ms_per_launch": graph_first[1]},
                  "later_replays_median": {"ms_per_step": graph_later[0] * 1e3,
                                           "events_ms_per_launch": graph_later[1]},
                  "replay_matches_setup_pass": graph_ok,
                  "what": f"the same {args.steps} launches captured once in a HIP graph, replayed after the same "
                          f"cache flush"},
        "kernel_events": {"ms_per_kernel": ext_ms, "min_ms": min(kernel_durations), "max_ms": max(kernel_durations),
                          "what": f"the timed region's {args.steps} launches queued behind a sleep kernel, each kernel's "
                                  f"own start and stop (hipExtLaunchKernel events; they end before the "
                                  f"end-of-kernel cache release)"},
        "tensor_launcher": {"ms_per_step": eager_wall * 1e3, "events_ms_per_launch": eager_ms,
                            "what": "one call per step of the tensor-level Python launcher "
                                    "(sapien_amd.native_fk_jac_rnea: device pointers looked up per call)"},
    }
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": (f"synthetic: seeded 7-dof chain (seed 0), q~U(limits), qd~N(0,1), qdd~N(0,var 4); step i reads "
                 f"buffer set i % {NSETS} ({min(args.steps, NSETS)} distinct sets in the timed region), "
                 f"Infinity Cache flushed before it"),
        "config": {"workload": "chain7 FK (8 link poses) + WORLD Jacobian (6x7) + RNEA, B=16384, FP32",
                   "envs": B, "dof": n, "links": L, "kernel": sa.kernel_variant(m),
                   "shape": ({"lanes_per_env": tuned[B][0], "joints_per_lane": tuned[B][1],
                              "envs_per_wavefront": 64 // tuned[B][0], "autotuned": not args.shape}
                             if tuned else None)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (traffic["FETCH_SIZE"]["bytes"] + traffic["WRITE_SIZE"]["bytes"]) if traffic else None,
                     "kernel_ms": kern_ms, "bytes_per_launch": BYTES_PER_EVAL * B,
                     "copy_gbs": {"value": copy_gbs, "what": "torch.neg over 256 MiB (read + write), 10 launches, "
                                                             "HIP events: the bandwidth an elementwise kernel reaches "
                                                             "on this GPU"}},
        "launch_modes": launch_modes,
    }
    if traffic and "VALU" in traffic:
        v = traffic["VALU"]
        # the SQ counts FLOPs per wavefront instruction (1 per add / mul /
        # transcendental, 2 per FMA; transcendentals included, _TRANS reports
        # them separately): x 64 lanes
        flops = 64 * v["SQ_INSTS_VALU_FLOPS_FP32"]
        tflops = flops / (kern_ms * 1e-3) / 1e12
        out["roofline"]["valu_fp32"] = {
            "flops_per_launch": flops, "achieved_tflops": tflops, "peak_tflops": FP32_PEAK_TFLOPS,
            "frac": tflops / FP32_PEAK_TFLOPS,
            "from_instruction_counts": 64 * (2 * v["SQ_INSTS_VALU_FMA_F32"] + v["SQ_INSTS_VALU_ADD_F32"] +
                                             v["SQ_INSTS_VALU_MUL_F32"] + v["SQ_INSTS_VALU_TRANS_F32"]),
            "valu_instructions_per_launch": v["SQ_INSTS_VALU"],
            "method": "rocprofv3 --pmc SQ_INSTS_VALU_FLOPS_FP32 x 64 lanes, cross-checked by "
                      "64 x (2 FMA + ADD + MUL + TRANS)_F32 instruction counts; SQ_INSTS_VALU; median of 16 "
                      "launches, over the launch duration kernel_ms"}
    if traffic:
        out["roofline"]["traffic_detail"] = {
            "fetch_bytes": traffic["FETCH_SIZE"]["bytes"], "write_bytes": traffic["WRITE_SIZE"]["bytes"],
            "fetch_scale": traffic["FETCH_SIZE"]["scale"], "write_scale": traffic["WRITE_SIZE"]["scale"],
            "method": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch, median of 16 launches over 16 "
                      "buffer sets), scaled so an elementwise kernel over 256 MiB reads and writes 256 MiB"}

    # the same kernel over other batch sizes (not the metric): how the launch
    # cost amortizes. Buffer sets rotate over > 600 MB so that inputs come
    # from HBM; autotuned shape per size; HIP events on the launch stream.
    sweep = []
    for Bs in (4096, 65536, 262144, 1048576):
        shp = tune.autotune(m, batch_sizes=(Bs,), use_cache=False)[Bs][:2] if tuned else None
        nset = max(2, -(-600_000_000 // (BYTES_PER_EVAL * Bs)))
        bsets = []
        for k in range(nset):
            qb = sets[k % NSETS]
            reps = -(-Bs // B)
            bsets.append([x.repeat(reps, 1)[:Bs].contiguous() for x in (qb["q"], qb["qd"], qb["qdd"])] +
                         [torch.empty(Bs, L, 7, device=dev), torch.empty(Bs, 6, n, device=dev),
                          torch.empty(Bs, n, device=dev)])
        for i in range(2 * nset):
            launch(Bs, *bsets[i % nset], stream.cuda_stream)
        reps_t = max(50, nset)
        # each kernel's own duration, the launches queued behind a sleep kernel
        kus = tune.kernel_us(lambda i, e0, e1: timed_launch(Bs, *bsets[i % nset], e0, e1, stream.cuda_stream), reps_t)
        gbs = BYTES_PER_EVAL * Bs / (kus * 1e-6) / 1e9
        sweep.append({"envs": Bs, "kernel_us": kus, "evals_s": Bs / (kus * 1e-6), "achieved_gbs": gbs,
                      "frac": gbs / HBM_PEAK_GBS, "shape": list(shp) if shp else None, "buffer_sets": nset})
        del bsets
    if tuned:
        tune.set_shape(m, B, tuned[B][0], tuned[B][1])
    out["batch_sweep"] = sweep

    # gradient throughput: 8192 environments of the 7-dof and 9-dof trees,
    # loss = |p_ee - FK(q*)|^2 + |tau_RNEA(q, qd, qdd) - tau*|^2. Forward-only
    # = the fused FK + RNEA evaluation of that loss's inputs; forward + backward
    # = the fused loss-and-gradient kernel (dq, dqd, dqdd).
    from sapien_amd.native import launcher_pose_torque_loss

    def graph_launch_s(fns, per_graph=50, replays=10):
        """Seconds per step of `fns` (each fn(stream) launches once) from HIP
        events around replays of a graph of per_graph steps: GPU time only,
        not the host's launch rate."""
        gcap = torch.cuda.Stream(dev)
        gcap.wait_stream(stream)
        with torch.cuda.stream(gcap):
            for fn in fns:  # warm-up outside capture
                fn(gcap.cuda_stream)
        stream.wait_stream(gcap)
        torch.cuda.synchronize()
        gg = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gg, stream=gcap):
            cst = torch.cuda.current_stream(dev).cuda_stream
            for _ in range(per_graph):
                for fn in fns:
                    fn(cst)
        stream.wait_stream(gcap)
        gg.replay()
        torch.cuda.synchronize()
        g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        g0.record(stream)
        for _ in range(replays):
            gg.replay()
        g1.record(stream)
        torch.cuda.synchronize()
        return g0.elapsed_time(g1) * 1e-3 / (per_graph * replays)

    grads = {}
    mixed = {}  # per tree: (forward, forward + backward) over the first 4096 environments
    for gname in ("chain7", "arm9"):
        grob = S.robot(gname)
        gm = m if gname == "chain7" else sa.PinocchioModel(grob.urdf, [0.0, 0.0, -9.81])
        if gname != "chain7":
            gm.set_joint_order(grob.joint_names)
            gm.set_link_order(grob.link_names)
        GB, gn, gL = 8192, grob.nv, grob.nlinks
        gq, gqd, gqdd, gts, gt = [torch.from_numpy(x).to(dev) for x in S.gradient_workload(grob, GB)]
        gt = gm.compute_forward_kinematics_batched(gt)[:, grob.ee_link, :3].contiguous()
        bufs = dict(P=torch.empty(GB, gL, 7, device=dev), T=torch.empty(GB, gn, device=dev),
                    loss=torch.empty(GB, device=dev), dq=torch.empty(GB, gn, device=dev),
                    dqd=torch.empty(GB, gn, device=dev), dqdd=torch.empty(GB, gn, device=dev))
        fwd = sa.native_fk_jac_rnea(gm)
        fb = launcher_pose_torque_loss(gm)
        if sa.kernel_variant(gm) == "lanes_per_joint":
            tune.autotune(gm, batch_sizes=(GB // 2, GB), use_cache=False)

        def run_f(st, nb=GB, fwd=fwd, gq=gq, gqd=gqd, gqdd=gqdd, bufs=bufs):
            fwd(nb, gq, gqd, gqdd, bufs["P"], None, bufs["T"], st)

        def run_fb(st, nb=GB, fb=fb, gq=gq, gqd=gqd, gqdd=gqdd, gts=gts, ee=grob.ee_link, gt=gt, bufs=bufs):
            fb(nb, gq, gqd, gqdd, gts, ee, gt, bufs["loss"], None, bufs["dq"], bufs["dqd"], bufs["dqdd"], st)

        mixed[gname] = (lambda st, f=run_f: f(st, GB // 2), lambda st, f=run_fb: f(st, GB // 2))
        res = {}
        for key, fn in (("forward_evals_s", run_f), ("fwd_bwd_evals_s", run_fb)):
            res[key] = GB / graph_launch_s([fn])
        res["ratio"] = res["fwd_bwd_evals_s"] / res["forward_evals_s"]
        res["kernel"] = sa.kernel_variant(gm)
        grads[gname] = res
    # one batch of 4096 + 4096 environments of the two trees: a forward (or a
    # loss + gradient) step launches both trees' kernels back to back
    res = {"envs": {"chain7": 4096, "arm9": 4096}}
    for key, idx in (("forward_evals_s", 0), ("fwd_bwd_evals_s", 1)):
        res[key] = 8192 / graph_launch_s([mixed[k][idx] for k in ("chain7", "arm9")])
    res["ratio"] = res["fwd_bwd_evals_s"] / res["forward_evals_s"]
    res["timing"] = ("HIP events around replays of a graph of 50 steps (each step the listed launches), 10 "
                     "replays after a warm-up one; per step = elapsed / 500")
    grads["mixed_chain7_arm9"] = res
    out["gradients"] = grads

    if not args.no_workloads:
        out["workloads"] = mixed_workloads(sa, S, np, torch, dev)

    # plain per-call drop-in path on the GPU (FP64, one configuration per call)
    q64, qd64, qdd64, _ = S.samples(rob, 256, 1001)
    t0 = time.perf_counter()
    for i in range(256):
        m.compute_forward_kinematics(q64[i])
        for l in range(L):
            m.get_link_pose(l)
        m.compute_full_jacobian(q64[i])
        m.compute_inverse_dynamics(q64[i], qd64[i], qdd64[i])
    plain = 256 / (time.perf_counter() - t0)
    out["plain_gpu"] = {"value": plain, "unit": "evals/s",
                        "sample": "256 x (compute_forward_kinematics + 8 get_link_pose + compute_full_jacobian + "
                                  "compute_inverse_dynamics), FP64 per call on the GPU"}
    out["speedup_vs_plain_gpu"] = value / plain

    if not args.no_cpu:
        from oracle.urdf_model import OracleModel

        o = OracleModel(rob.urdf)
        o.set_joint_order(rob.joint_names)
        o.set_link_order(rob.link_names)
        # host cores: the MI355X hosts this runs on allot 16 host cores to each
        # GPU, so the multi-core baseline uses at most 16 threads (and no more
        # than this process may run on)
        try:
            allowed = len(os.sched_getaffinity(0))
        except (AttributeError, OSError):
            allowed = os.cpu_count() or 1
        threads = max(1, min(16, allowed))
        cpu_model = ""
        try:
            with open("/proc/cpuinfo") as f:
                cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        # one seeded chunk of 1M samples, evaluated repeatedly until the sample
        # holds about cpu_seconds of CPU work
        chunk = 1 << 20
        q, qd, qdd, _ = S.samples(rob, chunk, 1001)
        q, qd, qdd = (np.ascontiguousarray(x) for x in (q, qd, qdd))
        bufs = o.fk_jac_rnea_timed(q[:1], qd[:1], qdd[:1], threads)[1:]
        bufs = tuple(np.zeros((chunk,) + b.shape[1:]) for b in bufs)
        o.fk_jac_rnea_timed(q, qd, qdd, threads, out=bufs)  # warm-up: touches every page once
        times = []  # per repeat, at least 5, until about cpu_seconds; the median repeat is the value
        while (sum(times) < args.cpu_seconds or len(times) < 5) and len(times) < 1000:
            times.append(o.fk_jac_rnea_timed(q, qd, qdd, threads, out=bufs)[0])
        secs, reps = sum(times), len(times)
        med = sorted(times)[reps // 2]
        out["cpu_baseline"] = {"value": chunk / med, "unit": "evals/s", "cores": threads, "kind": "port",
                               "repeats": reps, "min_max_evals_s": [chunk / max(times), chunk / min(times)],
                               "host": {"cpu_model": cpu_model, "os_cpu_count": os.cpu_count(),
                                        "affinity_cores": allowed,
                                        "threads_used": threads, "cap": "16 host cores per GPU, the host's allotment"},
                               "sample": f"{reps} x {chunk} seeded chain7 samples after a warm-up pass, FP64 FK + "
                                         f"8 link poses + WORLD Jacobian + RNEA per sample, C++ oracle statically "
                                         f"split over {threads} threads ({secs:.1f} s); value = the median repeat"}
        # one core: a quarter of the chunk until about a quarter of cpu_seconds
        q1, qd1, qdd1 = (x[: chunk // 4] for x in (q, qd, qdd))
        b1 = tuple(b[: chunk // 4] for b in bufs)
        t1 = []
        while (sum(t1) < args.cpu_seconds / 4 or len(t1) < 5) and len(t1) < 100:
            t1.append(o.fk_jac_rnea_timed(q1, qd1, qdd1, 1, out=b1)[0])
        s1, r1 = sum(t1), len(t1)
        out["cpu_baseline"]["single_core"] = {
            "value": (chunk // 4) / sorted(t1)[r1 // 2], "unit": "evals/s", "cores": 1, "repeats": r1,
            "sample": f"{r1} x {chunk // 4} samples, the same C++ oracle on one thread ({s1:.1f} s); value = the "
                      f"median repeat"}
        # the plain CPU reference as a user calls it: 1024 samples, one
        # configuration per call, FK + all link poses + WORLD Jacobian (FP64)
        qa = S.samples(rob, 1024, 1000)[0]
        t0 = time.perf_counter()
        for i in range(1024):
            fk = o.link_poses(qa[i:i + 1])[0]  # compute_forward_kinematics (every link pose cached)
            for l in range(L):  # get_link_pose(l): one call per link, a (p, q) pair each
                _ = (fk[l, :3].copy(), fk[l, 3:].copy())
            o.jacobian_world(qa[i:i + 1])  # compute_full_jacobian
        out["cpu_baseline"]["per_call_api"] = {
            "value": 1024 / (time.perf_counter() - t0), "unit": "evals/s", "cores": 1,
            "sample": "1024 seeded chain7 samples through the oracle's Python API, per sample one FK call, 8 "
                      "get_link_pose-style per-link reads and one full-Jacobian call (FP64), as the plain CPU "
                      "reference workload calls PinocchioModel"}
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
