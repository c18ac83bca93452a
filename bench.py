"""Benchmark: Mandelbulb distance-estimation throughput on one MI355X.

Headline workload (octree-adaptive sampling): root cube [-1.5, 1.5]^3, power 8,
15 iterations, bailout 2; the octree is refined level-synchronously on the GPU
to depth 8 (two forced levels, then split where |DE(centre)| < half diagonal
or NaN) and every surface leaf is sampled as an 8^3-cell span (9^3 corners,
the reference's one-cell overflow) with its crossing bits.  One step = the
whole refinement + span sampling.  value = distance evaluations per second
(centre + corner evaluations), inputs resident in HBM (they are generated on
the device from the root span), parity arithmetic (bit-exact loop state).

Prints one JSON line.  Options: --steps K --warmup W --mode parity|fast
--cpu-seconds S (budget of the CPU baseline sample; 0 disables it).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_F32 = 157.3e12   # MI355X dense FP32 vector peak (FLOP/s)
PEAK_HBM = 8.0e12     # MI355X HBM3E (B/s)
FLOP_PER_BODY = 75    # P8 loop body: magnitude 5, compare 1, dr 7, rotation 59, add 3
FLOP_BAILED = 6       # final magnitude + compare of a point that bails
FLOP_FINAL = 4        # ln, *r, *0.5, /dr


def algorithmic_flop(iters, max_iters: int) -> float:
    import torch
    it = iters.to(torch.int64)
    total = float(it.sum().item())
    bailed = float((iters < max_iters).sum().item())
    return FLOP_PER_BODY * total + FLOP_BAILED * bailed + FLOP_FINAL * iters.numel()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="parity", choices=["parity", "fast"])
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every workload shards by independent leaves, rays or jobs; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl

    c2 = wl.C2
    ctx = ca.Context(0, args.mode)
    shape = ca.Mandelbulb.classic(c2["max_iters"], c2["bailout"], ctx=ctx)
    hd = wl.half_diagonals(3.0, c2["max_depth"])
    res = c2["res"]
    per = (res + 1) ** 3

    # Size the buffers once (the refinement is deterministic).
    leaves, lc = dev.refine(shape, c2["root"], c2["min_depth"], c2["max_depth"], hd)
    nleaves = leaves.shape[0]
    leaves_buf = torch.empty((nleaves, 6), dtype=torch.float32, device="cuda")
    dists = torch.empty((nleaves, per), dtype=torch.float32, device="cuda")
    cross = torch.empty((nleaves, (res ** 3 + 63) // 64), dtype=torch.int64, device="cuda")
    center_evals = int(sum(int(lc[d]) for d in range(c2["min_depth"], c2["max_depth"] + 1)))
    corner_evals = nleaves * per
    evals = center_evals + corner_evals

    def step():
        lv, _ = dev.refine(shape, c2["root"], c2["min_depth"], c2["max_depth"], hd, leaves=leaves_buf)
        dev.sample_spans(shape, lv, res, out=dists, cross=cross)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = evals * args.steps / elapsed

    # Dominant kernel: the span-sampling refill kernel.  Its average launch
    # time is measured with HIP events on the stream it runs on (torch's
    # current stream), launching exactly what a step launches for it.
    reps = max(3, args.steps)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    dev.sample_spans(shape, leaves_buf, res, with_cross=False, out=dists)
    ev0.record()
    for _ in range(reps):
        dev.sample_spans(shape, leaves_buf, res, with_cross=False, out=dists)
    ev1.record()
    torch.cuda.synchronize()
    kernel_s = ev0.elapsed_time(ev1) / 1e3 / reps
    _, _, iters = dev.sample_spans(shape, leaves_buf, res, with_cross=False, with_iters=True)
    torch.cuda.synchronize()
    flop = algorithmic_flop(iters, c2["max_iters"])
    mean_iters = float(iters.to(torch.float64).mean().item())
    del iters
    achieved = flop / kernel_s
    traffic = None
    tf = ROOT / "profiles" / "spans_c2_traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {
        "bound": "valu",
        "achieved": achieved / 1e12,
        "peak": PEAK_F32 / 1e12,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_F32,
        "traffic": traffic,
        "kernel": "k_refill<P8Exact, SrcSpans, SinkDist>" if args.mode == "parity" else "k_refill<P8Fast, SrcSpans, SinkDist>",
        "kernel_ms": kernel_s * 1e3,
        "algorithmic_flop_per_launch": flop,
        "mean_loop_bodies": mean_iters,
        "hbm_write_gbs": corner_evals * 4 / kernel_s / 1e9,
    }

    # Reference CPU path (the oracle's op-exact restatement), timed on this
    # host on a bounded sample of the same leaves.
    cpu = None
    if args.cpu_seconds > 0:
        import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        lv_host = leaves_buf.cpu().numpy()
        rng = np.random.default_rng(0)
        s = O.mandelbulb(c2["max_iters"], c2["bailout"])
        m = min(256, nleaves)
        t = time.perf_counter()
        O.sample_spans(s, lv_host[rng.choice(nleaves, m, replace=False)], res, threads=threads)
        probe = time.perf_counter() - t
        m2 = int(min(nleaves, max(m, m * args.cpu_seconds / max(probe, 1e-3))))
        sample = lv_host[np.sort(rng.choice(nleaves, m2, replace=False))]
        t = time.perf_counter()
        O.sample_spans(s, sample, res, threads=threads)
        dt = time.perf_counter() - t
        cpu = {"value": m2 * per / dt, "unit": "evals/s", "cores": threads, "kind": "port",
               "sample": f"{m2} of {nleaves} depth-8 surface leaves (seeded), 9^3 corners each, "
                         f"oracle restatement, {threads} threads, {dt:.1f} s"}

    line = {
        "metric": "Mandelbulb DE evals/s (p=8,15 it) + 1080p rays/s; fwd+adjoint vs fwd+6-tap FD",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "octree-adaptive span sampling: root [-1.5,1.5]^3, P8, 15 it, bailout 2, "
                        "refine to depth 8 (|DE(centre)| < half diagonal), 8^3-cell spans per surface leaf",
            "mode": args.mode,
            "surface_leaves": nleaves,
            "center_evals": center_evals,
            "corner_evals": corner_evals,
            "level_counts": [int(x) for x in lc],
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
