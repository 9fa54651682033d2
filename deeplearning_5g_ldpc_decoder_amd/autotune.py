"""Autotuner for the packed-tile mapping of algo-1 batches.

For each (base graph, lifting size) it times decode launches with k
codeblocks per group, k in {1, 2, 4, ..., as many as fit}, on a batch large
enough to keep every CU busy (tiles >= 2 per CU), and keeps the k with the
highest codeblock throughput; for the lifting sizes the lane-per-edge schedule
takes (one wave per codeblock, a layer's edges one per lane: Zc <= 3 for BG1,
<= 6 for BG2) it times that schedule too and records the faster as the
class's schedule.  Then it times the tile schedule on a
mixed-size batch: one workgroup per tile against a persistent grid that pulls
the tiles from a device queue, and keeps the faster.  The table goes to
tile_table.json beside this file; BatchPlan applies it (k through
ldpc_mi355_set_tile_codeblocks, the schedule as its default) before building
plans.

    python -m deeplearning_5g_ldpc_decoder_amd.autotune [--zc 2 3 ...] [--bg 1 2]
    python -m deeplearning_5g_ldpc_decoder_amd.autotune --schedule-only
    python -m deeplearning_5g_ldpc_decoder_amd.autotune --edge-lanes-only
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib
from .workloads import LIFTING_SIZES, payload_for, waterfall_esn0_db

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tile_table.json")
# lifting sizes the lane-per-edge schedule takes (degree * Zc <= 64 lanes)
EDGE_LANES_MAX_ZC = {1: 3, 2: 6}
ROW = [1, 8, 22936, 0, 1280, 32767, -32767]


def _time_plan(torch, plan, d_llr, reps):
    out = plan.alloc_outputs("cuda")
    s = torch.cuda.current_stream()
    for _ in range(2):
        plan.launch(d_llr, out, stream=s.cuda_stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(s)
    for _ in range(reps):
        plan.launch(d_llr, out, stream=s.cuda_stream)
    ev[1].record(s)
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / 1e3 / reps


def tune_class(torch, bgn, zc, reps=4, distinct=16, cus=256):
    from .batch import BatchPlan, synth_batch
    crc = 2
    A = payload_for(bgn, zc, crc)
    if A is None:
        crc = 0
        A = payload_for(bgn, zc, crc)
    if A is None:
        return None
    B = A + (16 if crc == 0 else 24)
    E = (2 * B) // 2 * 2                      # rate 1/2, QPSK
    cfg = (bgn, A, crc, E, 0, 2, 0)
    llr, _ = synth_batch(cfg, distinct, -(waterfall_esn0_db(B / E, 2, B) + 0.5), seed=zc + 97 * bgn)
    d_llr = torch.from_numpy(llr.reshape(-1)).cuda()
    L = _lib.load()
    cap = min(384 // zc, 192)
    cands = sorted({k for k in (1, 2, 4, 8, 16, 32, 64, 128, cap) if k <= cap})
    res = {}
    for k in cands:
        L.ldpc_mi355_set_tile_codeblocks(bgn, zc, k)
        n = max(2 * cus * k, 512)
        offs = [(i % distinct) * E for i in range(n)]
        plan = BatchPlan([cfg] * n, ROW, llr_offsets=offs, flags=BatchPlan.PACK,
                         schedule="oneshot")
        dt = _time_plan(torch, plan, d_llr, reps)
        plan.close()
        res[k] = n / dt
    L.ldpc_mi355_set_tile_codeblocks(bgn, zc, 0)
    best = max(res, key=res.get)
    out = dict(k=int(best), payload=A, crc=crc, E=E, schedule="tiles",
               codeblocks_per_s={str(k): float(v) for k, v in res.items()})
    if zc <= EDGE_LANES_MAX_ZC[bgn]:
        # the lane-per-edge schedule: one wave per codeblock, a layer's edges
        # one per lane
        n = 8 * cus
        offs = [(i % distinct) * E for i in range(n)]
        plan = BatchPlan([cfg] * n, ROW, llr_offsets=offs,
                         flags=BatchPlan.NO_PACK | BatchPlan.EDGE_LANES, schedule="oneshot")
        dt = _time_plan(torch, plan, d_llr, reps)
        plan.close()
        out["edge_lanes_codeblocks_per_s"] = n / dt
        if n / dt > res[best]:
            out["schedule"] = "edge_lanes"
    return out


def tune_schedule(torch, n=4096, reps=20):
    """One-shot (a workgroup per tile) against persistent (a resident grid
    pulling tiles from a device queue) on a mixed-size batch of n codeblocks,
    with the k table already applied; outputs must agree."""
    from .batch import BatchPlan
    from .workloads import mixed_batch
    cfgs, llrs, _ = mixed_batch(n, seed=5)
    d_llr = torch.from_numpy(np.concatenate(llrs)).cuda()
    res, outs = {}, {}
    for name in ("oneshot", "persistent"):
        plan = BatchPlan(cfgs, ROW, flags=BatchPlan.PACK, schedule=name)
        res[name] = _time_plan(torch, plan, d_llr, reps)
        outs[name] = {k: v.cpu().numpy() for k, v in plan.decode(d_llr).items()}
        plan.close()
    same = all(np.array_equal(outs["oneshot"][k], outs["persistent"][k]) for k in outs["oneshot"])
    if not same:
        raise RuntimeError("persistent and one-shot tile launches disagree")
    return dict(persistent=bool(res["persistent"] < res["oneshot"]), codeblocks=n,
                oneshot_ms=res["oneshot"] * 1e3, persistent_ms=res["persistent"] * 1e3)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--bg", type=int, nargs="*", default=[1, 2])
    ap.add_argument("--zc", type=int, nargs="*", default=list(LIFTING_SIZES))
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--schedule-only", action="store_true",
                    help="keep the k table in --out and re-time only the tile schedule")
    ap.add_argument("--edge-lanes-only", action="store_true",
                    help="keep the table in --out and re-time only the classes the "
                         "lane-per-edge schedule takes")
    args = ap.parse_args(argv)
    import torch
    assert torch.cuda.is_available(), "autotuning needs the GPU"
    if args.edge_lanes_only:
        from .batch import apply_tuning
        with open(args.out) as f:
            table = json.load(f)
        apply_tuning(args.out)
        for bgn in (1, 2):
            for zc in LIFTING_SIZES:
                if zc > EDGE_LANES_MAX_ZC[bgn]:
                    continue
                r = tune_class(torch, bgn, zc)
                if r is None:
                    continue
                table["classes"][f"{bgn}:{zc}"] = r
                print(f"autotune: BG{bgn} Zc={zc}: k={r['k']} "
                      f"({r['codeblocks_per_s'][str(r['k'])]:.3g} cb/s), lane-per-edge "
                      f"{r['edge_lanes_codeblocks_per_s']:.3g} cb/s -> {r['schedule']}",
                      flush=True)
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1)
        return 0
    if args.schedule_only:
        from .batch import apply_tuning
        with open(args.out) as f:
            table = json.load(f)
        apply_tuning(args.out)
        table["schedule"] = tune_schedule(torch)
        print("autotune: schedule", table["schedule"], flush=True)
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1)
        return 0
    table = {"row": ROW, "device": torch.cuda.get_device_name(0), "classes": {}}
    t0 = time.time()
    for bgn in args.bg:
        for zc in args.zc:
            r = tune_class(torch, bgn, zc)
            if r is None:
                continue
            table["classes"][f"{bgn}:{zc}"] = r
            print(f"autotune: BG{bgn} Zc={zc}: k={r['k']} "
                  f"({r['codeblocks_per_s'][str(r['k'])]:.3g} cb/s)", flush=True)
    table["seconds"] = time.time() - t0
    from .batch import apply_tuning
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    apply_tuning(args.out)
    table["schedule"] = tune_schedule(torch)
    print("autotune: schedule", table["schedule"], flush=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print("autotune: wrote", args.out, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
