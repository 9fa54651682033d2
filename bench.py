#!/usr/bin/env python3
"""Benchmark: decoded info Gbps of the 5G NR LDPC decode path on one MI355X.

Workload (one "step" = one batched decode launch over the resident batch):
BG1 Zc=384 (K=8448), A=8424 + CRC24B, rv0, E=25344 (R~1/3), QPSK over AWGN at
Es/N0 = 1 dB, LLR x128 quantized to int16 (clip +-10000), 256 codeblocks,
decoder row [1, 8, 22936, 0, 1280, 32767, -32767] (int16 layered min-sum,
8 iterations max, CRC24B early stop).  Inputs are synthesized on the host,
copied to HBM once, and stay resident.

Two timed regions of K steps each:
- single stream: K launches back to back on one stream; HIP events around
  each launch give the kernel time behind `roofline` (the same launches
  `--profile` runs under rocprofv3);
- pipelined (`value`): K steps over --inflight streams (default 2) with one
  output set per stream, so a launch's codeblocks start on the CUs the
  previous launch's early-stopping codeblocks free (one workgroup holds a
  CU; a launch otherwise ends with its slowest codeblock).  Every step still
  decodes the whole batch; its outputs are checked equal to the
  single-stream ones.

Prints ONE JSON line.  Also reported: host-visible batch latency p50/p99, the
per-codeblock ABI path (c_ldpc_5g_decode, one launch per codeblock) and the
CPU baseline (the oracle's -O3 restatement of the reference decoder on the
host cores).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CFG = (1, 8424, 2, 25344, 0, 2, 0)
ROW = [1, 8, 22936, 0, 1280, 32767, -32767]
BATCH = 256
NOISE_DB = -1.0        # Es/N0 = 1 dB with unit-energy symbols
EDGES_BG1 = 316
HBM_PEAK = 8.0e12
CLOCK_HZ = 2.4e9       # MI355X peak engine clock
SIMDS = 256 * 4        # 256 CUs x 4 SIMDs
# VALU issue: a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles
VALU_PEAK = SIMDS * CLOCK_HZ / 4    # wave-instructions per second
# LDS: 128 B per clock per CU (CDNA3 figure, taken for gfx950) x 256 CUs x 2.4 GHz
LDS_PEAK = 128 * 256 * CLOCK_HZ


def algorithmic_bytes(iters, E=25344, B=8448, zc=384, w=8):
    """Per launch: sum over codeblocks of 2E + B/8 + I_exec * w * nnz * Zc
    (LQ and Lr read + written once per edge per iteration, int16)."""
    iters = np.asarray(iters, np.int64)
    return int(np.sum(2 * E + B // 8 + iters * w * EDGES_BG1 * zc))


def lds_bytes(iters, zc=384):
    """Belief traffic through LDS per launch: every edge reads and writes its
    int16 LQ entry once per executed iteration (messages stay in VGPRs)."""
    iters = np.asarray(iters, np.int64)
    return int(np.sum(iters * 4 * EDGES_BG1 * zc))


def load_profile(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def load_traffic():
    return load_profile("hbm_traffic_minsum_i16_bg1_zc384_b256.json").get(
        "traffic_bytes_per_launch")


def valu_instructions(iters):
    """Wave-VALU instructions of one launch of the specialized BG1 Zc=384
    kernel: per codeblock (its 6 waves) a fixed part (rate recovery, CRC
    setup, outputs) plus a per-iteration part, each a deterministic
    instruction stream
    (counted with SQ_INSTS_VALU at two iteration limits,
    profiles/valu_counts_minsum_i16_bg1_zc384_b256.json), times the
    iterations this launch executed."""
    prof = load_profile("valu_counts_minsum_i16_bg1_zc384_b256.json")
    if not prof:
        return None
    iters = np.asarray(iters, np.int64)
    return (len(iters) * prof["per_codeblock_wave_instr"] +
            int(iters.sum()) * prof["per_iteration_wave_instr"])


def host_cpu():
    """CPU model, logical CPUs visible to this process and physical cores."""
    model, phys = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            pid, core = None, None
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    core = v
                elif not k and pid is not None:
                    phys.add((pid, core))
                    pid = core = None
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "logical_cpus_visible": visible,
            "logical_cpus_host": os.cpu_count(), "physical_cores_host": len(phys) or None}


def cpu_baseline(llr, threads):
    """Oracle restatement of the reference algo-1 decoder (-O3, scalar loop
    structure of c_lib) on the host cores, one codeblock per task."""
    from oracle import oracle as O
    L = O.lib()
    n = llr.shape[0]
    hard = np.zeros((n, 1056), np.uint8)
    st = np.zeros(n, np.int32)
    its = np.zeros(n, np.int32)
    d = O.dec(ROW)
    secs = L.orc_decode_batch_i16(O._p(llr), O._p(hard), O._p(st), O._p(its), n, *CFG[:6],
                                  ctypes.byref(d), threads)
    return secs, st, its, hard


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def bench_mixed(args):
    """Mixed-size batch in one launch (BG1/BG2, all 51 lifting sizes, rate
    0.2-0.9, Qm 2-8, Es/N0 within 1 dB of each rate's waterfall), packed
    tiles with the autotuned mapping, launched one workgroup per tile or as a
    persistent grid pulling tiles from a device queue, vs one codeblock per
    workgroup.  `value` is the faster tile schedule (single stream) with
    --inflight launches in flight (as the headline's pipelined region;
    outputs checked equal), beside its single-stream rate."""
    import torch
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    from deeplearning_5g_ldpc_decoder_amd.workloads import mixed_batch
    n = args.batch
    log(f"synthesizing {n} mixed codeblocks")
    cfgs, llrs, pays = mixed_batch(n, seed=3)
    d_llr = torch.from_numpy(np.concatenate(llrs)).cuda()
    info = sum(c[1] for c in cfgs)
    res = {}
    for name, flags, sched in (("packed", BatchPlan.PACK, "oneshot"),
                               ("persistent", BatchPlan.PACK, "persistent"),
                               ("plain", BatchPlan.NO_PACK, "oneshot")):
        plan = BatchPlan(cfgs, ROW, flags=flags, schedule=sched)
        out = plan.alloc_outputs("cuda")
        s = torch.cuda.current_stream()
        for _ in range(args.warmup):
            plan.launch(d_llr, out, stream=s.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(s)
        for _ in range(args.steps):
            plan.launch(d_llr, out, stream=s.cuda_stream)
        e1.record(s)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = out["status"].cpu().numpy()
        hard = plan.split_hard(out["hard"].cpu().numpy())
        good = sum(int(np.array_equal(np.unpackbits(h)[:c[1]], p))
                   for h, c, p in zip(hard, cfgs, pays))
        res[name] = {"gbps": info * args.steps / wall / 1e9,
                     "kernel_ms": e0.elapsed_time(e1) / args.steps, "tiles": plan.tiles,
                     "bler": float(np.mean(st != 0)), "payload_correct": good,
                     "mean_iterations": float(out["iters"].float().mean().item())}
        if name == "packed":
            ref_out = {f: out[f].cpu().numpy() for f in ("status", "iters", "hard")}
        elif name == "persistent":
            res[name]["outputs_equal_packed"] = all(
                np.array_equal(out[f].cpu().numpy(), ref_out[f]) for f in ref_out)
        if name in ("packed", "persistent") and args.inflight > 1:
            # the same launches with `inflight` in flight on their own streams
            # and outputs (as the headline's pipelined region)
            k = args.inflight
            pstreams = [torch.cuda.Stream() for _ in range(k)]
            pouts = [plan.alloc_outputs("cuda") for _ in range(k)]
            for i in range(max(args.warmup, k)):
                plan.launch(d_llr, pouts[i % k], stream=pstreams[i % k].cuda_stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i in range(args.steps):
                plan.launch(d_llr, pouts[i % k], stream=pstreams[i % k].cuda_stream)
            torch.cuda.synchronize()
            pwall = time.perf_counter() - t0
            same = all(np.array_equal(po[f].cpu().numpy(), out[f].cpu().numpy())
                       for po in pouts for f in ("status", "iters", "hard"))
            res[name]["pipelined"] = {"inflight": k, "gbps": info * args.steps / pwall / 1e9,
                                      "ms_per_step": pwall * 1e3 / args.steps,
                                      "outputs_equal_single_stream": same}
        if name == "packed":
            res[name]["graph"] = mixed_graph_leg(plan, d_llr, out, info, args)
        plan.close()
        log(f"{name}: {res[name]}")
    # the schedule with the faster single-stream launch carries `value`
    best = max(("packed", "persistent"), key=lambda k: res[k]["gbps"])
    pipe = res[best].get("pipelined")
    line = {"metric": "decoded info Gbps, mixed-size batch in one launch", "value":
            pipe["gbps"] if pipe else res[best]["gbps"], "unit": "Gbps", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": pipe["ms_per_step"] if pipe else res[best]["kernel_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int16",
            "data": "synthetic",
            "config": {"workload": f"mixed_bg12_allzc_minsum_i16_8it_batch{n}",
                       "global_batch": n, "decoder_row": ROW, "parallelism": "single-gpu",
                       "schedule": best},
            "packed": res["packed"], "persistent": res["persistent"],
            "one_codeblock_per_workgroup": res["plain"],
            "packing_speedup": res[best]["gbps"] / res["plain"]["gbps"]}
    print(json.dumps(line))


def mixed_graph_leg(plan, d_llr, ref, info, args):
    """The packed plan's launches captured in one HIP graph (a graph of
    `per_graph` back-to-back launches, replayed until --steps launches ran),
    timed with events on the replay stream; every replay's outputs are the
    eager ones."""
    import torch
    per_graph = 10
    s = torch.cuda.Stream()
    gout = plan.alloc_outputs("cuda")
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for _ in range(per_graph):
            plan.launch(d_llr, gout, stream=s.cuda_stream)
    reps = max(1, args.steps // per_graph)
    same = True
    for _ in range(2):          # replays checked against the eager outputs
        for f in ("status", "iters", "hard"):
            gout[f].zero_()
        g.replay()
        torch.cuda.synchronize()
        same &= all(np.array_equal(gout[f].cpu().numpy(), ref[f].cpu().numpy())
                    for f in ("status", "iters", "hard"))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(s):
        e0.record(s)
        for _ in range(reps):
            g.replay()
        e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / (reps * per_graph)
    return {"launches_per_graph": per_graph, "replays": reps, "ms_per_launch": ms,
            "gbps": info / (ms * 1e-3) / 1e9, "outputs_equal_eager": bool(same)}


def bench_train(args):
    """Training steps of the trainable decoder: BG1 Zc=96 (A=2088 + CRC24B,
    E=6336, R=1/3, QPSK), 10 unrolled iterations, alpha and beta per
    (iteration, row degree), clip_llr = 32767/128, batch 2048 with Es/N0 uniform
    in [0, 2] dB per codeblock, BCE on -LLR of the B bits, float64 grad-norm
    clip, Adam.  A step = forward + loss + backward + optimizer step over one
    fresh batch synthesized on the device inside the step (encode, rate
    match, channel, float rate recovery)."""
    import torch
    from deeplearning_5g_ldpc_decoder_amd.trainer import BG1_ZC96_WORKLOAD, _clip_grad_norm_float64
    from deeplearning_5g_ldpc_decoder_amd.unrolled import (CTableNeuralLayeredMinSumDecoder,
                                                           load_ldpc_table_config)
    w = BG1_ZC96_WORKLOAD
    n = args.batch if args.batch != BATCH else w["batch_size"]
    tc = load_ldpc_table_config(w["bgn"], w["A"], w["crctype"], w["E"], w["rv"], w["Qm"])
    clip = 32767.0 / 128.0
    T = w["num_iterations"]
    model = CTableNeuralLayeredMinSumDecoder(
        tc, num_iterations=T, alpha_mode=w["alpha_mode"], beta_mode=w["beta_mode"],
        init_alpha=w["init_alpha"], init_beta=w["init_beta"], learnable_beta=True,
        clip_llr=clip).cuda()
    opt = torch.optim.Adam(model.parameters(), lr=w["lr"])
    crit = torch.nn.BCEWithLogitsLoss()
    B = tc.B
    stream = torch.cuda.current_stream()

    from deeplearning_5g_ldpc_decoder_amd.unrolled import synth_float_batch_device
    nz_rng = np.random.default_rng(1)

    def step(k, ev=None):
        if ev:
            ev[3].record(stream)
        # fresh inputs every step, synthesized on the device
        x, tgt = synth_float_batch_device(tc, nz_rng.uniform(-2.0, 0.0, n), seed=1000 + k)
        opt.zero_grad(set_to_none=True)
        if ev:
            ev[0].record(stream)
        _, q = model(x, return_llr=True)
        loss = crit(-q[:, :B], tgt)
        if ev:
            ev[1].record(stream)
        loss.backward()
        if ev:
            ev[2].record(stream)
        _clip_grad_norm_float64(model.parameters(), 1.0)
        opt.step()
        return loss

    log("warmup")
    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    losses = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        losses.append(step(k, evs[k]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    synth_ms = float(np.mean([e[3].elapsed_time(e[0]) for e in evs]))
    fwd_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    bwd_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    losses = [float(v) for v in losses]
    ms = wall * 1e3 / args.steps
    log(f"kernels: {ms:.3f} ms/step (synthesis {synth_ms:.3f} ms, forward {fwd_ms:.3f} ms, "
        f"backward {bwd_ms:.3f} ms)")

    k = model.kernel()
    line = {"metric": "training codeblocks/s (BG1 Zc=96, 10 unrolled iterations, batch 2048)",
            "value": n / ms * 1e3, "unit": "codeblocks/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "train_bg1_zc96_minsum_f32_10it_batch2048",
                       "global_batch": n, "iterations": T, "alpha_mode": w["alpha_mode"],
                       "beta_mode": w["beta_mode"], "clip_llr": clip,
                       "parallelism": "single-gpu"},
            "input_synthesis_ms": synth_ms, "forward_ms": fwd_ms, "backward_ms": bwd_ms,
            "record_MB": k.record_bytes(n) / 2 ** 20, "scratch_MB": k.scratch_bytes(n) / 2 ** 20,
            "loss_first": losses[0], "loss_last": losses[-1]}
    print(json.dumps(line))


def bench_serve(args):
    """Served decode loop: requests of 1-64 mixed-size codeblocks (base graph,
    lifting size, rate 0.2-0.9, Qm 2-8, Es/N0 within 1 dB of each rate's
    waterfall) arrive as a Poisson process from 16 UEs x 4 HARQ processes,
    each with a 0.5 ms deadline; failed transmissions come back after 1 ms with
    the next redundancy version (0, 2, 3, 1) and soft-combine into device
    HARQ buffers.  Continuous batching over --serve-contexts launch contexts
    (one HIP stream each), the loop in C++ (ldpc_mi355_serve_run; --serve-loop
    python runs the Python one).  Request LLRs are resident in HBM and decoded
    in place (--serve-llr host_mapped: in pinned host memory, read in place
    over PCIe; host_copy: copied from there into a 256 MB device pool that
    admission budgets).  First the unloaded latency (each template
    alone), the capacity (requests/s with every request queued at once), then
    offered loads of 0.2-1.0 x capacity."""
    # HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues per process;
    # streams sharing a queue run their kernels one after another, so with
    # fewer queues than contexts concurrent batches serialize.  Set before
    # the first HIP call of the process.
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.serve_hw_queues)
    from deeplearning_5g_ldpc_decoder_amd.serving import (DevicePool, GpuExecutor, HarqSlots,
                                                          NativeServer, RequestPool, Server,
                                                          poisson_arrivals)
    log("synthesizing 128 request templates (4 redundancy versions each)")
    pool = RequestPool(128, seed=11, max_cbs=64, threads=min(16, os.cpu_count() or 8))
    harq = HarqSlots(16, 4, 64)
    llr_pool = DevicePool(256 << 20)
    native = args.serve_loop == "native"
    if args.serve_contexts is None:
        args.serve_contexts = 4 if native else 2
    rng = np.random.default_rng(5)
    if native:
        srv = NativeServer(pool, harq, llr_pool, n_contexts=args.serve_contexts,
                           max_batch_codeblocks=args.serve_batch, deadline_s=0.5e-3,
                           harq_rtt_s=1e-3,
                           llr_source=args.serve_llr)
    else:
        ex = GpuExecutor(pool, harq, llr_pool, n_contexts=args.serve_contexts,
                         max_batch_codeblocks=args.serve_batch)
        srv = Server(ex, deadline_s=0.5e-3, harq_rtt_s=1e-3)

    def run(arr, tpl=None, ues=None):
        n = len(arr)
        tpl = rng.integers(0, len(pool), n) if tpl is None else tpl
        ues = rng.integers(0, 16, n) if ues is None else ues
        return srv.run(arr, tpl, ues=ues)

    run(np.zeros(64))          # warmup
    # unloaded latency: every template alone in the system (arrivals 3 ms
    # apart, longer than a request and its retransmission take)
    st = run(np.arange(len(pool)) * 3e-3, np.arange(len(pool)), np.zeros(len(pool), np.int64))
    alone = np.array([r.latency for r in st.requests if r.rv_index == 0]) * 1e3
    unloaded = {"p50_ms": float(np.percentile(alone, 50)),
                "p99_ms": float(np.percentile(alone, 99)), "requests": len(alone)}
    log(f"unloaded latency p50 {unloaded['p50_ms']:.3f} ms p99 {unloaded['p99_ms']:.3f} ms")
    nsat = 600
    st = run(np.zeros(nsat))
    sat = st.summary(0.5e-3)
    cap = sum(1 for r in st.requests if r.rv_index == 0) / st.wall_s
    log(f"capacity {cap:.0f} requests/s, {sat['info_gbps']:.2f} Gbps saturated")
    sweep = []
    for f in (0.2, 0.4, 0.6, 0.8, 1.0):
        rate = f * cap
        dur = min(1.5, 1500.0 / rate)
        s = run(poisson_arrivals(rate, dur, seed=int(f * 10))).summary(0.5e-3)
        s.update(load=f, offered_requests_per_s=rate, duration_s=dur)
        log(f"load {f:.1f}: p50 {s['latency_ms']['p50']:.3f} ms p99 {s['latency_ms']['p99']:.3f} "
            f"ms, miss {s['deadline_miss_rate']:.3f}, {s['info_gbps']:.2f} Gbps; from arrival "
            f"p99 {s['latency_from_arrival_ms']['p99']:.3f} ms, miss "
            f"{s['deadline_miss_rate_from_arrival']:.3f}")
        sweep.append(s)
    if not native:
        ex.close()
    batch_path = serve_mix_batched(pool, args)
    served_cbs = sat["codeblocks"] / st.wall_s
    log(f"batch path, the same mix: {batch_path['codeblocks_per_s']:.0f} codeblocks/s, "
        f"{batch_path['info_gbps']:.2f} Gbps decoded; served at saturation {served_cbs:.0f} "
        "codeblocks/s")
    mid = sweep[2]
    line = {"metric": "served decode p99 latency at 0.6x capacity (mixed sizes, HARQ, 0.5 ms "
                      "deadline)",
            "value": mid["latency_ms"]["p99"], "unit": "ms", "n_gpus": 1,
            "steps": sum(s["requests"] for s in sweep), "warmup": 64,
            "ms_per_step": mid["latency_ms"]["mean"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic",
            "config": {"workload": "serve_mixed_1to64cb_harq16x4_poisson",
                       "decoder_row": ROW, "contexts": args.serve_contexts,
                       "loop": args.serve_loop,
                       "llr_source": args.serve_llr if native else "host_copy",
                       "hw_queues": args.serve_hw_queues,
                       "max_batch_codeblocks": args.serve_batch,
                       "llr_pool_MB": 256, "harq_slots": 16 * 4 * 64,
                       "parallelism": "single-gpu"},
            "capacity_requests_per_s": cap, "unloaded": unloaded, "saturated": sat,
            "saturated_codeblocks_per_s": served_cbs, "batch_path": batch_path,
            "served_vs_batch_codeblock_rate": served_cbs / batch_path["codeblocks_per_s"],
            "latency_note": "latency_ms runs from a transmission joining the decode queue (new "
                            "data once its UE has a free HARQ process, a retransmission when "
                            "due) to its results on the host; latency_from_arrival_ms also "
                            "counts new data's wait for a free HARQ process",
            "value_from_arrival": mid["latency_from_arrival_ms"]["p99"],
            "sweep": sweep}
    print(json.dumps(line))


def serve_mix_batched(pool, args):
    """The served request mix through the batch path: the first transmission
    (rv 0) of every request template, all codeblocks in one launch of the
    same decoder row, timed over --steps launches with HIP events.  Gives the
    rate the GPU decodes this mix at when nothing limits concurrency (no
    deadlines, HARQ processes or per-batch round trips)."""
    import torch
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    cfgs, parts, info = [], [], []
    for t in range(len(pool)):
        parts.append(np.asarray(pool.llr(t, 0)[0]))
        cfgs += [tuple(c[:4]) + (0, c[5], 0) for c in pool.cfgs[t]]
        info.append(pool.cb_info[t])
    info = np.concatenate(info)
    plan = BatchPlan(cfgs, ROW)
    d_llr = torch.from_numpy(np.concatenate(parts)).cuda()
    out = plan.alloc_outputs("cuda")
    s = torch.cuda.current_stream()
    for _ in range(3):
        plan.launch(d_llr, out, stream=s.cuda_stream)
    steps = max(10, min(args.steps, 200))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        plan.launch(d_llr, out, stream=s.cuda_stream)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    ok = out["status"].cpu().numpy()[:len(cfgs)] == 0
    plan.close()
    return {"codeblocks": len(cfgs), "launch_ms": ms, "steps": steps,
            "codeblocks_per_s": len(cfgs) / (ms * 1e-3),
            "info_gbps": float(info.sum()) / (ms * 1e-3) / 1e9,
            "passed_info_gbps": float(info[ok].sum()) / (ms * 1e-3) / 1e9,
            "codeblock_bler": float(1.0 - ok.mean())}


BLER_ROWS = {4: [4, 20, 22936, 0, 1280, 32767, -32767],   # FP32 min-sum
             0: [0, 20, 1, 0, 1280, 20, -32768]}           # FP32 BP


def bench_bler(args):
    """FP32 min-sum (algo 4; --bler-algo 0: FP32 BP), 20 iterations, CRC early
    stop, BLER sweep of the BG2 workload: Zc 104 / 208 (B = 1040 / 2080), R 0.2 / 0.5, Qm 2 / 4 (Gray
    QAM, max-log LLRs), 8 SNR points per config spanning its waterfall
    (Es/N0 from 2.5 dB below to 1.5 dB above the waterfall model of
    workloads.py; reported as the harness's snr_power_db, noise_power_db =
    -snr - 10 log10 Qm), args.batch codeblocks
    per point (default 4096) synthesized on the GPU (ldpc_mi355_synth_device),
    each point one batched launch; BLER with its 99 % Wilson interval, and
    decode throughput."""
    import torch
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan, synth_device
    from deeplearning_5g_ldpc_decoder_amd.ldpc_sim import bler_confidence_interval
    from deeplearning_5g_ldpc_decoder_amd.workloads import waterfall_esn0_db
    row = BLER_ROWS[args.bler_algo]
    name = {4: "FP32 min-sum", 0: "FP32 BP"}[args.bler_algo]
    tag = {4: "fp32ms", 0: "fp32bp"}[args.bler_algo]
    n = args.batch if args.batch != BATCH else 4096
    npts = 8
    table = []
    gpu_s, info = 0.0, 0
    for A in (1016, 2056):
        for R in (0.2, 0.5):
            for qm in (2, 4):
                B = A + 24
                E = int(round(B / R / qm)) * qm
                cfg = (2, A, 2, E, 0, qm, 0)
                wf = waterfall_esn0_db(B / E, qm, B)
                snrs = np.linspace(wf - 2.5, wf + 1.5, npts) - 10 * np.log10(qm)
                plan = BatchPlan([cfg] * n, row)
                out = plan.alloc_outputs("cuda")
                pts = []
                for si, snr in enumerate(snrs):
                    d = synth_device(cfg, n, float(-snr - 10 * np.log10(qm)),
                                     seed=1000 * si + qm)["llr"].reshape(-1)
                    plan.launch(d, out)
                    torch.cuda.synchronize()
                    t = time.perf_counter()
                    plan.launch(d, out)
                    torch.cuda.synchronize()
                    gpu_s += time.perf_counter() - t
                    info += n * A
                    st = out["status"].cpu().numpy()
                    e = int(np.count_nonzero(st != 0))
                    lo, hi = bler_confidence_interval(e, n)
                    pts.append({"snr_db": float(snr), "bler": e / n, "ci99": [lo, hi],
                                "mean_iterations": float(out["iters"].float().mean())})
                plan.close()
                table.append({"cfg": cfg, "rate": R, "points": pts})
                log(f"BG2 A={A} R={R} Qm={qm}: " +
                    " ".join(f"{p['bler']:.3f}" for p in pts))
    line = {"metric": f"{name} decoded info Gbps over the BG2 BLER sweep",
            "value": info / gpu_s / 1e9, "unit": "Gbps", "n_gpus": 1,
            "steps": len(table) * npts, "warmup": 1,
            "ms_per_step": gpu_s * 1e3 / (len(table) * npts), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": f"bler_bg2_zc104_208_{tag}_20it_batch{n}",
                       "decoder_row": row, "parallelism": "single-gpu"},
            "sweep": table}
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--latency-reps", type=int, default=1000)
    ap.add_argument("--max-iterations", type=int, default=None,
                    help="with --profile only: decoder row iteration limit (to count the "
                         "per-iteration and fixed VALU instructions of the kernel)")
    ap.add_argument("--inflight", type=int, default=2,
                    help="headline workload: launches in flight in the pipelined timed "
                         "region (one stream and output set each; 1 = single stream)")
    ap.add_argument("--profile", action="store_true",
                    help="batch launches only (for rocprofv3): no latency loop, no "
                         "per-codeblock path, no CPU leg")
    ap.add_argument("--workload", default="bg1_zc384",
                    choices=["bg1_zc384", "mixed", "train", "serve", "bler"],
                    help="bg1_zc384: the headline workload (default); mixed: 8192 mixed-size "
                         "codeblocks in one launch (use --batch to change); train: training "
                         "steps of the trainable decoder (BG1 Zc=96, batch 2048); serve: the "
                         "served loop (Poisson arrivals, HARQ, continuous batching); bler: the "
                         "FP32 BG2 BLER sweep")
    ap.add_argument("--bler-algo", type=int, default=4, choices=[4, 0],
                    help="bler workload: FP32 min-sum (4) or FP32 BP (0)")
    ap.add_argument("--serve-batch", type=int, default=256,
                    help="serve workload: codeblocks per batch at most")
    ap.add_argument("--serve-contexts", type=int, default=None,
                    help="serve workload: launch contexts (one stream each) in flight at most "
                         "(default 4 with the native loop, 2 with the Python loop)")
    ap.add_argument("--serve-hw-queues", type=int, default=8,
                    help="serve workload: GPU_MAX_HW_QUEUES for the process (hardware queues "
                         "the HIP streams are spread over; at most 32)")
    ap.add_argument("--serve-llr", default="device",
                    choices=["device", "host_mapped", "host_copy"],
                    help="serve workload, native loop: request LLRs resident in HBM (as a GPU "
                         "demapper leaves them) and decoded in place; in pinned host memory, "
                         "decoded in place over PCIe; or in pinned host memory and copied into "
                         "the batch's region of the device pool")
    ap.add_argument("--serve-loop", default="native", choices=["native", "python"],
                    help="serve workload: the C++ loop (ldpc_mi355_serve_run) or the Python one")
    args = ap.parse_args()
    if args.profile:
        args.no_cpu = True
    if args.max_iterations is not None and not args.profile:
        ap.error("--max-iterations changes the workload: only with --profile")
    if not 1 <= args.serve_hw_queues <= 32:
        ap.error("--serve-hw-queues: 1 to 32")
    if args.gpus != 1:
        print(f"bench.py: the decoder runs on one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard codeblocks by index, with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "mixed":
        if args.batch == BATCH:
            args.batch = 8192
        return bench_mixed(args)
    if args.workload == "train":
        return bench_train(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "bler":
        return bench_bler(args)

    import torch
    from deeplearning_5g_ldpc_decoder_amd import _lib
    from deeplearning_5g_ldpc_decoder_amd._lib import DecoderConfig, LdpcInConfig, ptr
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan, synth_batch

    n = args.batch
    log("synthesizing inputs")
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    llr, payload = synth_batch(CFG, n, NOISE_DB, seed=1, threads=os.cpu_count() or 8)
    d_llr = torch.from_numpy(llr.reshape(-1)).to(dev)
    row = list(ROW)
    if args.max_iterations is not None:
        row[1] = args.max_iterations
    plan = BatchPlan([CFG] * n, row)
    out = plan.alloc_outputs(dev)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream

    log("warmup")
    for _ in range(args.warmup):
        plan.launch(d_llr, out, stream=sh)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        plan.launch(d_llr, out, stream=sh)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])

    status = out["status"].cpu().numpy()
    iters = out["iters"].cpu().numpy()
    hard = out["hard"].cpu().numpy().reshape(n, 1056)
    decoded = np.unpackbits(hard, axis=1)[:, :CFG[1]]
    good = int(np.sum(np.all(decoded == payload, axis=1)))

    info_bits = n * CFG[1]
    single_gbps = info_bits * args.steps / wall / 1e9
    log(f"single stream: timed {args.steps} steps in {wall:.3f} s")

    # pipelined region: launches alternate over `inflight` non-blocking
    # streams (default priority), each with its own outputs
    inflight = 1 if args.profile else max(1, args.inflight)
    pipe_same = None
    if inflight > 1:
        pstreams = [torch.cuda.Stream(dev) for _ in range(inflight)]
        pouts = [plan.alloc_outputs(dev) for _ in range(inflight)]
        for k in range(max(args.warmup, inflight)):
            plan.launch(d_llr, pouts[k % inflight], stream=pstreams[k % inflight].cuda_stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            plan.launch(d_llr, pouts[k % inflight], stream=pstreams[k % inflight].cuda_stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        log(f"pipelined ({inflight} in flight): timed {args.steps} steps in {wall:.3f} s")
        pipe_same = all(np.array_equal(po["status"].cpu().numpy(), status) and
                        np.array_equal(po["iters"].cpu().numpy(), iters) and
                        np.array_equal(po["hard"].cpu().numpy().reshape(n, 1056), hard)
                        for po in pouts)
    gbps = info_bits * args.steps / wall / 1e9
    ms_per_step = wall * 1e3 / args.steps
    # host-visible latency of one batch (launch -> results synchronized)
    lat = []
    for _ in range(0 if args.profile else args.latency_reps):
        t = time.perf_counter()
        plan.launch(d_llr, out, stream=sh)
        stream.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat if lat else [float("nan")])

    # per-codeblock reference ABI path: one launch + copies per codeblock
    per_cb_gbps, per_cb = None, None
    if not args.profile:
        L = _lib.load()
        ic = LdpcInConfig(*CFG)
        dc = DecoderConfig(*ROW)
        hb = np.zeros(1056, np.uint8)
        n_pc = 256
        for i in range(32):
            L.c_ldpc_5g_decode(ptr(llr[i % n]), ptr(hb), ctypes.byref(ic), ctypes.byref(dc))
        calls = []
        for i in range(n_pc):
            t = time.perf_counter()
            L.c_ldpc_5g_decode(ptr(llr[i % n]), ptr(hb), ctypes.byref(ic), ctypes.byref(dc))
            calls.append(time.perf_counter() - t)
        calls = np.array(calls)
        per_cb_gbps = n_pc * CFG[1] / calls.sum() / 1e9
        per_cb = {"calls": n_pc, "us_mean": float(calls.mean() * 1e6),
                  "us_p50": float(np.median(calls) * 1e6), "us_max": float(calls.max() * 1e6),
                  "call_of_max": int(np.argmax(calls)),
                  "gbps_p50": CFG[1] / float(np.median(calls)) / 1e9}

    avg_kernel_s = float(kernel_ms.mean()) / 1e3
    # the binding resource is VALU issue (decoder state stays in LDS and
    # VGPRs; HBM carries little more than the compulsory I/O)
    vi = valu_instructions(iters)
    traffic = load_traffic()
    alg = algorithmic_bytes(iters)
    roof = {"bound": "valu",
            "achieved": vi / avg_kernel_s / 1e9 if vi else None,
            "peak": VALU_PEAK / 1e9, "unit": "G wave-instr/s",
            "frac": vi / avg_kernel_s / VALU_PEAK if vi else None,
            "traffic": traffic,
            "valu_wave_instr_per_launch": vi,
            "kernel_ms_avg": float(kernel_ms.mean()),
            "note": "achieved = counted wave-VALU instructions of this launch (fixed part per "
                    "codeblock + per-iteration part x executed iterations) / kernel time; peak = "
                    "1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction. One codeblock's 6 "
                    "waves sit 2-2-1-1 on a CU's 4 SIMDs, so full issue on the busiest SIMDs is "
                    "frac 0.75",
            "hbm": {"achieved": traffic / avg_kernel_s / 1e9 if traffic else None,
                    "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": traffic / avg_kernel_s / HBM_PEAK if traffic else None,
                    "compulsory_bytes_per_launch": int(n * (2 * 25344 + 1056))},
            "reference_layout_effective": {
                "bytes_per_launch": alg, "GBps": alg / avg_kernel_s / 1e9,
                "note": "the reference's per-edge LQ/Lr traffic (8 B per edge per iteration) "
                        "over kernel time; those bytes never reach HBM here"},
            "lds": {"achieved": lds_bytes(iters) / avg_kernel_s / 1e9,
                    "peak": LDS_PEAK / 1e9, "unit": "GB/s",
                    "frac": lds_bytes(iters) / avg_kernel_s / LDS_PEAK}}
    if vi and inflight > 1:
        # informational: the same instruction count per step over the
        # pipelined region's time per step, where the next launch's
        # workgroups fill the CUs that early-stopping codeblocks free
        roof["pipelined"] = {
            "achieved": vi / (ms_per_step / 1e3) / 1e9,
            "frac": vi / (ms_per_step / 1e3) / VALU_PEAK,
            "note": "wave-VALU instructions per step / pipelined ms_per_step (the step's "
                    "decode overlapped with the previous one); not a per-launch figure"}

    log("per-codeblock path done")
    cpu = None
    if not args.no_cpu:
        info = host_cpu()
        # one GPU's share of the host: at most 16 worker threads on this box
        threads = max(1, min(info["logical_cpus_visible"] or 1, 16))
        log(f"cpu baseline: {threads} threads ({info})")
        reps = 3
        secs = 0.0
        for _ in range(reps):
            s_, cst, cits, chard = cpu_baseline(llr, threads)
            secs += s_
        same = bool(np.array_equal(cst, status) and np.array_equal(cits, iters) and
                    np.array_equal(chard, hard))
        n1 = 16
        s1, _, _, _ = cpu_baseline(llr[:n1], 1)
        single_mbps = n1 * CFG[1] / s1 / 1e6
        cpu = {"value": reps * info_bits / secs / 1e9, "unit": "Gbps", "cores": threads,
               "kind": "port",
               "sample": f"{reps} passes over the same {n}-codeblock batch on {threads} threads "
                         f"(one codeblock per task, {reps * info_bits / secs / 1e6:.2f} Mbit/s; "
                         f"bit-exact with the GPU result: {same}); single core: {n1} codeblocks "
                         f"on 1 thread, {single_mbps:.3f} Mbit/s",
               "single_core_mbps": single_mbps,
               "all_threads_mbps": reps * info_bits / secs / 1e6,
               "mean_iterations": float(cits.mean()),
               "cpu_model": info["model"], "logical_cpus_visible": info["logical_cpus_visible"],
               "logical_cpus_host": info["logical_cpus_host"],
               "physical_cores_host": info["physical_cores_host"],
               "threads_note": "one GPU's share of the host cores is 16 threads on this box"}

    line = {
        "metric": "decoded info Gbps (BG1 Zc=384, 8 it) + p99 decode latency, BLER-matched",
        "value": gbps, "unit": "Gbps", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic",
        "config": {"workload": "bg1_zc384_minsum_i16_8it_batch256", "global_batch": n,
                   "codeblock": {"bgn": 1, "A": 8424, "crc": "24B", "E": 25344, "Qm": 2,
                                 "rv": 0, "EsN0_dB": 1.0},
                   "decoder_row": ROW, "parallelism": "single-gpu"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)),
                       "p99": float(np.percentile(lat, 99))},
        "pipeline": {"inflight": inflight, "single_stream_gbps": single_gbps,
                     "outputs_equal_single_stream": pipe_same},
        "mean_iterations": float(iters.mean()), "max_iterations_executed": int(iters.max()),
        "bler": float(np.mean(status != 0)), "payload_correct": good,
        "per_codeblock_abi_gbps": per_cb_gbps,
        "per_codeblock_abi": per_cb,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    plan.close()


if __name__ == "__main__":
    main()
