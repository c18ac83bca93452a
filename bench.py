"""Benchmark: Mandelbulb distance-estimation throughput on one MI355X.

Headline workload (octree-adaptive sampling + mesh extraction): root cube
[-1.5, 1.5]^3, power 8, 15 iterations, bailout 2; the octree is refined
level-synchronously on the GPU to depth 8 (two forced levels, then split where
|DE(centre)| < half diagonal or NaN), every surface leaf is sampled as an
8^3-cell span (9^3 corners, the reference's one-cell overflow), and naive
surface nets turns each leaf into vertices (6-tap difference normals, 7
distance evaluations per vertex, as the reference) and quads.  One step = the
whole refinement + span sampling + mesh extraction.  value = distance
evaluations per second (centre + corner + vertex evaluations), inputs resident
in HBM (they are generated on the device from the root span), parity
arithmetic (bit-exact loop state).

Prints one JSON line.  Options: --steps K --warmup W --mode parity|fast
--cpu-seconds S (budget of the CPU baseline sample; 0 disables it).
"""
from __future__ import annotations

import argparse
import json
import re
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
FLOP_DEAD = 62        # rotation + add of the last body of an evaluation at the iteration cap:
                      # the rotated point is never read, so it is not counted
FLOP_VERTEX = 19      # 6-tap normal per vertex: 6 stencil adds, 3 differences, normalize 10
FLOP_STEP = 11        # sphere-tracing step: pos += dir d (6), t += d, eps + eps_rel t (2), 2 compares
FLOP_RAY = 33         # pinhole ray: u 6, v 5, direction 12, normalize 10


def de_flop(bodies: int, bails: int, evals: int) -> float:
    """Algorithmic FLOP of P8 distance evaluations: 75 per loop body, of which
    only 13 (magnitude, compare, dr) in the last body of an evaluation that
    reaches the iteration cap (every evaluation that does not bail), 6 per
    evaluation that bails (its last magnitude and compare), 4 per final formula."""
    capped = evals - bails
    return FLOP_PER_BODY * bodies - FLOP_DEAD * capped + FLOP_BAILED * bails + FLOP_FINAL * evals


def count_work(ctx, fn):
    """(loop bodies, bailed evaluations, evaluations) that the refill launches
    issued by fn() ran, from the library's work counters."""
    import torch
    cnt = torch.zeros(3, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    ctx.count_work(cnt)
    try:
        fn()
        torch.cuda.synchronize()
    finally:
        ctx.count_work(None)
    return tuple(int(v) for v in cnt.cpu())


def event_ms(fn, reps, before=None):
    """Average duration of fn() over reps launches, each bracketed by HIP
    events on torch's current stream (the stream the kernels run on);
    before() (untimed) runs ahead of each launch."""
    import torch
    total = 0.0
    for _ in range(reps):
        if before is not None:
            before()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        total += e0.elapsed_time(e1)
    return total / reps


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="parity", choices=["parity", "fast"])
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here "
              "(every workload shards by independent leaves, rays or jobs).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl

    from cantucci_amd import autotune

    oc = wl.OCTREE
    ctx = ca.Context(0, args.mode)
    tuned = autotune.apply(ctx, "spans")
    tuned_mesh = autotune.apply(ctx, "mesh")
    shape = ca.Mandelbulb.classic(oc["max_iters"], oc["bailout"], ctx=ctx)
    hd = wl.half_diagonals(3.0, oc["max_depth"])
    res = oc["res"]
    per = (res + 1) ** 3

    # Size the buffers once (the refinement is deterministic).
    leaves, lc = dev.refine(shape, oc["root"], oc["min_depth"], oc["max_depth"], hd)
    nleaves = leaves.shape[0]
    leaves_buf = torch.empty((nleaves, 6), dtype=torch.float32, device="cuda")
    dists = torch.empty((nleaves, per), dtype=torch.float32, device="cuda")
    cross = torch.empty((nleaves, (res ** 3 + 63) // 64), dtype=torch.int64, device="cuda")
    center_evals = int(sum(int(lc[d]) for d in range(oc["min_depth"], oc["max_depth"] + 1)))
    corner_evals = nleaves * per
    mesh_out = {}

    def step():
        lv, _ = dev.refine(shape, oc["root"], oc["min_depth"], oc["max_depth"], hd, leaves=leaves_buf)
        dev.sample_spans(shape, lv, res, with_cross=False, out=dists)
        # the mesh's count pass classifies the cells and writes the crossing words
        mesh_out["m"] = dev.mesh(shape, lv, dists, res, "fd6", cross=cross)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    voff, ioff, _, _ = mesh_out["m"]
    nverts = int(voff[-1].item())
    nquads = int(ioff[-1].item()) // 6
    del mesh_out["m"]
    vertex_evals = 7 * nverts
    evals = center_evals + corner_evals + vertex_evals
    ms_per_step = 1e3 * elapsed / args.steps
    value = evals * args.steps / elapsed
    phases = _phases(dev, torch, shape, oc, hd, leaves_buf, dists, cross, res)

    # The step's two large kernels, each timed alone with HIP events on the
    # stream it runs on (torch's current stream), launching exactly what a
    # step launches for it; their algorithmic FLOP come from the library's
    # work counters on one such launch.
    #   span sampling: k_refill<P8Exact, SrcSpans, SinkDist> (sample_spans)
    #   vertex taps:   k_mesh_taps<P8Exact> (the vertex pass of the mesh)
    reps = max(3, args.steps)
    dev.sample_spans(shape, leaves_buf, res, with_cross=False, out=dists)
    span_ms = event_ms(lambda: dev.sample_spans(shape, leaves_buf, res, with_cross=False, out=dists), reps)
    span_w = count_work(ctx, lambda: dev.sample_spans(shape, leaves_buf, res, with_cross=False, out=dists))
    span_flop = de_flop(*span_w)
    # records as the mesh's emit pass leaves them for the vertex pass
    _, _, vdef, _ = dev.mesh(shape, leaves_buf, dists, res, "deferred", cross=cross)
    vdef = vdef.clone()
    vwork = torch.empty_like(vdef)
    taps_ms = event_ms(lambda: dev.mesh_vertices(shape, vwork, "fd6"), reps, before=lambda: vwork.copy_(vdef))
    vwork.copy_(vdef)
    taps_w = count_work(ctx, lambda: dev.mesh_vertices(shape, vwork, "fd6"))
    taps_flop = de_flop(*taps_w) + FLOP_VERTEX * nverts
    del vdef, vwork
    traffic = {}
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # summaries (copy-kernel calibrated).
    for key, fname in (("spans", "spans_octree_traffic.md"), ("taps", "mesh_taps_octree_traffic.md")):
        tf = ROOT / "profiles" / fname
        traffic[key] = None
        if tf.exists():
            m = re.search(r"^bytes_per_launch:\s*(\d+)", tf.read_text(), re.M)
            traffic[key] = int(m.group(1)) if m else None
    suffix = "Exact" if args.mode == "parity" else "Fast"
    kernels = {
        "taps": {
            "kernel": f"k_mesh_taps<P8{suffix}>", "kernel_ms": taps_ms,
            "algorithmic_flop_per_launch": taps_flop, "loop_bodies": taps_w[0], "bailed_evals": taps_w[1],
            "evals": taps_w[2], "vertices": nverts,
            "achieved": taps_flop / (taps_ms / 1e3) / 1e12, "frac": taps_flop / (taps_ms / 1e3) / PEAK_F32,
            "traffic": traffic["taps"],
            # position and half-steps read, the whole 28-byte record written
            "algorithmic_bytes": nverts * (24 + 28),
        },
        "spans": {
            "kernel": f"k_refill<P8{suffix}, SrcSpans, SinkDist>", "kernel_ms": span_ms,
            "algorithmic_flop_per_launch": span_flop, "loop_bodies": span_w[0], "bailed_evals": span_w[1],
            "evals": span_w[2], "mean_loop_bodies": span_w[0] / max(1, span_w[2]),
            "achieved": span_flop / (span_ms / 1e3) / 1e12, "frac": span_flop / (span_ms / 1e3) / PEAK_F32,
            "traffic": traffic["spans"], "algorithmic_bytes": corner_evals * 4,
            "hbm_write_gbs": corner_evals * 4 / (span_ms / 1e3) / 1e9,
        },
    }
    dom = max(kernels, key=lambda k: kernels[k]["kernel_ms"])
    kd = kernels[dom]
    roofline = {
        "bound": "valu",
        "achieved": kd["achieved"],
        "peak": PEAK_F32 / 1e12,
        "unit": "TFLOP/s",
        "frac": kd["frac"],
        "traffic": kd["traffic"],
        "kernel": kd["kernel"],
        "kernel_ms": kd["kernel_ms"],
        "algorithmic_flop_per_launch": kd["algorithmic_flop_per_launch"],
        "kernels": kernels,
    }

    # Reference CPU path (the oracle's op-exact restatement of the whole
    # naive_surface_nets per leaf), timed on this host on a bounded sample of
    # the same leaves, threaded over leaves like the reference's pool.
    cpu = None
    if args.cpu_seconds > 0:
        import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        lv_host = leaves_buf.cpu().numpy()
        rng = np.random.default_rng(0)
        s = O.mandelbulb(oc["max_iters"], oc["bailout"])
        m = min(256, nleaves)
        t = time.perf_counter()
        O.mesh_batch(s, lv_host[rng.choice(nleaves, m, replace=False)], res, threads=threads)
        probe = time.perf_counter() - t
        m2 = int(min(nleaves, max(m, m * args.cpu_seconds / max(probe, 1e-3))))
        sample = lv_host[np.sort(rng.choice(nleaves, m2, replace=False))]
        t = time.perf_counter()
        vo, _, _, _ = O.mesh_batch(s, sample, res, threads=threads)
        dt = time.perf_counter() - t
        cpu_evals = m2 * per + 7 * int(vo[-1])
        one = sample[:2000]
        t = time.perf_counter()
        vo1, _, _, _ = O.mesh_batch(s, one, res, threads=1)
        dt1 = time.perf_counter() - t
        cpu = {"value": cpu_evals / dt, "unit": "evals/s", "cores": threads, "kind": "port",
               "sample": f"{m2} of {nleaves} depth-8 surface leaves (seeded): 9^3 corners + 7 per vertex, "
                         f"whole surface nets, oracle restatement, {threads} threads, {dt:.1f} s",
               "value_1thread": (len(one) * per + 7 * int(vo1[-1])) / dt1,
               "sample_1thread": f"the first {len(one)} leaves of that sample on one thread, {dt1:.2f} s",
               "host": _host_cpu()}
        # The reference application's own meshing workload on the same path.
        ref_spans = _reference_default_spans()
        t = time.perf_counter()
        O.mesh_batch(O.mandelbulb(6, 2.5), ref_spans, 64, threads=threads)
        ref_s = time.perf_counter() - t
        t = time.perf_counter()
        O.mesh_batch(O.mandelbulb(6, 2.5), ref_spans[:4], 64, threads=1)
        ref1_s = (time.perf_counter() - t) * len(ref_spans) / 4
        cpu["reference_default_mesh"] = {
            "workload": "64 leaves of [-1.2,1.2]^3 at res 64, classic(6, 2.5), whole surface nets",
            "ms": ref_s * 1e3, "threads": threads, "ms_1thread_extrapolated": ref1_s * 1e3,
            "sample_1thread": "4 of the 64 leaves on one thread, scaled to 64"}

    extra = {}
    if args.mode == "parity":
        try:
            extra["contract_mode"] = bench_contract(args, oc, hd, leaves_buf, dists, nverts, nquads)
        except Exception as e:  # a secondary leg must not hide the headline
            extra["contract_mode"] = {"error": f"{type(e).__name__}: {e}"}
    for name, fn in (("grid_64", bench_grid), ("trace_1080p", bench_trace), ("normals_shell", bench_normals),
                     ("mixed_jobs", bench_jobs), ("reference_default_mesh", bench_ref_default),
                     ("point_query", bench_point_query)):
        try:
            extra[name] = fn(ctx, args)
        except Exception as e:  # a secondary leg must not hide the headline
            extra[name] = {"error": f"{type(e).__name__}: {e}"}
    ref_leg = extra.get("reference_default_mesh", {})
    if cpu is not None and "gpu_ms" in ref_leg:
        ref_leg["cpu_ms"] = cpu["reference_default_mesh"]["ms"]
        ref_leg["speedup_vs_cpu"] = ref_leg["cpu_ms"] / ref_leg["gpu_ms"]

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
            "workload": "octree-adaptive sampling + mesh extraction: root [-1.5,1.5]^3, P8, 15 it, bailout 2, "
                        "refine to depth 8 (|DE(centre)| < half diagonal), 8^3-cell spans per surface leaf, "
                        "naive surface nets with 6-tap normals",
            "mode": args.mode,
            "surface_leaves": nleaves,
            "center_evals": center_evals,
            "corner_evals": corner_evals,
            "vertex_evals": vertex_evals,
            "vertices": nverts,
            "quads": nquads,
            "phases_ms": phases,
            "level_counts": [int(x) for x in lc],
            "launch": tuned,
            "launch_mesh": tuned_mesh,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


def bench_contract(args, oc, hd, leaves_buf, parity_dists, nverts_parity, nquads_parity):
    """The headline step with FMA contraction (CTH_MODE_CONTRACT: the parity
    sequence with products fused into following sums, IEEE division and
    square roots): step time, both kernels' rooflines, and the distance of its
    corner distances from the parity mode's on every corner (parity is the
    reference's operation sequence; the FP64 comparison is the GPU test
    tests/test_gpu_contract.py)."""
    import torch

    import cantucci_amd as ca
    from cantucci_amd import autotune
    from cantucci_amd import device as dev

    ctx = ca.Context(0, "contract")
    autotune.apply(ctx, "spans")
    autotune.apply(ctx, "mesh")
    shape = ca.Mandelbulb.classic(oc["max_iters"], oc["bailout"], ctx=ctx)
    res = oc["res"]
    nleaves = leaves_buf.shape[0]
    # this mode's own refinement can split a few centres differently
    lv, lc = dev.refine(shape, oc["root"], oc["min_depth"], oc["max_depth"], hd)
    nleaves_contract = lv.shape[0]
    center_evals = int(sum(int(lc[d]) for d in range(oc["min_depth"], oc["max_depth"] + 1)))
    lv_buf = torch.empty((max(nleaves, nleaves_contract), 6), device="cuda")
    dists = torch.empty((lv_buf.shape[0], (res + 1) ** 3), device="cuda")
    cross = torch.empty((lv_buf.shape[0], (res ** 3 + 63) // 64), dtype=torch.int64, device="cuda")
    del lv
    out = {}

    def step():
        lv, _ = dev.refine(shape, oc["root"], oc["min_depth"], oc["max_depth"], hd, leaves=lv_buf)
        m = lv.shape[0]
        dev.sample_spans(shape, lv, res, with_cross=False, out=dists[:m])
        out["m"] = dev.mesh(shape, lv, dists[:m], res, "fd6", cross=cross[:m])

    step()
    torch.cuda.synchronize()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t) / reps
    voff, ioff, _, _ = out.pop("m")
    nverts, nquads = int(voff[-1].item()), int(ioff[-1].item()) // 6
    evals = center_evals + nleaves_contract * (res + 1) ** 3 + 7 * nverts
    lv = lv_buf[:nleaves_contract]
    dv = dists[:nleaves_contract]
    span_ms = event_ms(lambda: dev.sample_spans(shape, lv, res, with_cross=False, out=dv), 3)
    span_flop = de_flop(*count_work(ctx, lambda: dev.sample_spans(shape, lv, res, with_cross=False, out=dv)))
    _, _, vdef, _ = dev.mesh(shape, lv, dv, res, "deferred", cross=cross[:nleaves_contract])
    vdef = vdef.clone()
    vwork = torch.empty_like(vdef)
    taps_ms = event_ms(lambda: dev.mesh_vertices(shape, vwork, "fd6"), 3, before=lambda: vwork.copy_(vdef))
    vwork.copy_(vdef)
    taps_flop = de_flop(*count_work(ctx, lambda: dev.mesh_vertices(shape, vwork, "fd6"))) + FLOP_VERTEX * nverts
    del vdef, vwork
    # the contraction mode's corner distances on the parity run's leaves
    dp = parity_dists.ravel()
    dc = dists[:nleaves]
    dev.sample_spans(shape, leaves_buf, res, with_cross=False, out=dc)
    dc = dc.ravel()
    fin = torch.isfinite(dc) & torch.isfinite(dp)
    rel = (dc[fin].double() - dp[fin].double()).abs() / dp[fin].double().abs().clamp_min(1e-30)
    tail = float((rel > 1e-5).double().mean().item())
    flips = float((torch.signbit(dc) != torch.signbit(dp)).double().mean().item())
    del rel, fin
    ctx.close()
    return {"workload": "the octree step in CTH_MODE_CONTRACT (FMA contraction, IEEE div/sqrt)",
            "value": evals / (ms / 1e3), "ms_per_step": ms, "vertices": nverts, "quads": nquads,
            "vertices_parity": nverts_parity, "quads_parity": nquads_parity,
            "taps": {"kernel_ms": taps_ms, "algorithmic_flop_per_launch": taps_flop,
                     "frac": taps_flop / (taps_ms / 1e3) / PEAK_F32},
            "spans": {"kernel_ms": span_ms, "algorithmic_flop_per_launch": span_flop,
                      "frac": span_flop / (span_ms / 1e3) / PEAK_F32},
            "surface_leaves": nleaves_contract, "surface_leaves_parity": nleaves,
            "corners_rel_gt_1e-5_vs_parity": tail, "corner_sign_differs_from_parity": flips}


def _host_cpu() -> str:
    """CPU model and logical CPU count of the host the baseline ran on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs (threads used: see cores)"


def _phases(dev, torch, shape, oc, hd, leaves_buf, dists, cross, res, reps=3):
    """Per-phase times of a step, the median of `reps` steps (HIP events on
    torch's current stream)."""
    runs = {"refine": [], "sample_spans": [], "mesh": []}
    for _ in range(reps):
        marks = []

        def mark():
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            marks.append(e)

        torch.cuda.synchronize()
        mark()
        lv, _ = dev.refine(shape, oc["root"], oc["min_depth"], oc["max_depth"], hd, leaves=leaves_buf)
        mark()
        dev.sample_spans(shape, lv, res, with_cross=False, out=dists)
        mark()
        m = dev.mesh(shape, lv, dists, res, "fd6", cross=cross)
        mark()
        torch.cuda.synchronize()
        for name, a, b in (("refine", 0, 1), ("sample_spans", 1, 2), ("mesh", 2, 3)):
            runs[name].append(marks[a].elapsed_time(marks[b]))
        del m
    return {name: sorted(v)[len(v) // 2] for name, v in runs.items()}


def _time_ms(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t) / reps


def bench_grid(ctx, args):
    """64^3 grid over [-1.5,1.5]^3 (divisor 63), P8 / 15 / 2: one launch is a
    few microseconds of work, so 100 launches are captured in a HIP graph, and
    the graph is replayed 10 times (1000 launches)."""
    import ctypes as C

    import torch

    import cantucci_amd as ca
    from cantucci_amd import autotune
    from cantucci_amd import device as dev
    from cantucci_amd._lib import check, lib

    autotune.apply(ctx, "grid")
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    n = 64
    out = torch.empty(n ** 3, device="cuda")
    s = torch.cuda.Stream()
    st = (C.c_float * 3)(-1.5, -1.5, -1.5)
    en = (C.c_float * 3)(1.5, 1.5, 1.5)
    launches = 100

    def body():
        for _ in range(launches):
            check(lib.cth_de_grid_dev(ctx.handle, C.byref(shape.desc), st, en, n, 63.0, out.data_ptr(), None,
                                      s.cuda_stream))

    g = dev.Graph(body, s)
    g.launch()
    s.synchronize()
    reps = 10  # 1000 launches
    t = time.perf_counter()
    for _ in range(reps):
        g.launch()
    s.synchronize()
    per_launch = (time.perf_counter() - t) / (reps * launches)
    single = _time_ms(lambda: dev.de_grid(shape, (-1.5,) * 3, (1.5,) * 3, n, 63.0), 20)
    g.close()
    # The same grid scaled up (divisor m - 1), one eager launch each.
    scaled = {}
    for m in (256, 512):
        ms = _time_ms(lambda: dev.de_grid(shape, (-1.5,) * 3, (1.5,) * 3, m, float(m - 1)), 5)
        scaled[f"{m}^3"] = {"evals_per_s": m ** 3 / (ms / 1e3), "ms": ms}
    # Other P8 settings on the 64^3 grid: the reference application's
    # classic(6, 2.5) and classic(8, 5.0) (max_iters, bailout).
    variants = {}
    for mi, bo in ((6, 2.5), (8, 5.0)):
        other = ca.Mandelbulb.classic(mi, bo, ctx=ctx)
        ms = _time_ms(lambda: dev.de_grid(other, (-1.5,) * 3, (1.5,) * 3, n, 63.0), 20)
        variants[f"classic({mi}, {bo})"] = {"us_per_launch_eager": ms * 1e3}
    return {"workload": "64^3 grid, P8, 15 it, bailout 2", "evals_per_s": n ** 3 / per_launch,
            "us_per_launch_graph": per_launch * 1e6, "us_per_launch_eager": single * 1e3,
            "scaled": scaled, "other_settings_64^3": variants}


def bench_point_query(ctx, args):
    """Latency of one host-buffer query, the reference's per-frame calls:
    the fly camera's min_distance_from (src/control/fly.rs:91-92) and the 25
    focus rays of get_focii (src/mesh/mod.rs:205-243)."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev  # noqa: F401  (torch loaded first)

    shape = ca.Mandelbulb.classic(6, 2.5, ctx=ctx)
    p = (-3.0, 0.1, 0.2)
    for _ in range(20):
        shape.min_distance_from(p)
    reps = 2000
    t = time.perf_counter()
    for _ in range(reps):
        shape.min_distance_from(p)
    one = (time.perf_counter() - t) / reps
    from cantucci_amd.mesh import get_focii

    # the app's projection (src/app.rs:98): fov 1 rad, near 4e-5, far 10
    def focii():
        get_focii(shape, (-3.0, 0.0, 0.0), (1.0, 0.0, 0.0), 1.0, 16 / 9, 0.00004, 10.0, 5)

    for _ in range(20):
        focii()
    t = time.perf_counter()
    for _ in range(500):
        focii()
    trace = (time.perf_counter() - t) / 500
    return {"workload": "classic(6, 2.5): one min_distance_from; get_focii's 25 focus rays built and marched "
                        "on the device (eps 1e-6, 100 steps)",
            "us_min_distance_from": one * 1e6, "us_focus_rays": trace * 1e6}


def bench_ref_default(ctx, args):
    """The reference application's own meshing workload: Mandelbulb::classic(6,
    2.5), the bounding box [-1.2, 1.2]^3 pre-split twice into 64 leaves
    (src/mesh/mod.rs:52-56) and every leaf meshed by naive surface nets at
    RESOLUTION = 64 (src/mesh/mod.rs:133): 64 x 65^3 corners + 7 evaluations
    per vertex.  GPU: one batched sample + mesh (the CPU time of the same
    workload is taken in the cpu_baseline leg)."""
    import torch

    import cantucci_amd as ca
    from cantucci_amd import device as dev

    shape = ca.Mandelbulb.classic(6, 2.5, ctx=ctx)
    spans = _reference_default_spans()
    res = 64
    sp = torch.from_numpy(spans).cuda()
    out = torch.empty((len(spans), (res + 1) ** 3), device="cuda")

    def run():
        dev.sample_spans(shape, sp, res, with_cross=False, out=out)
        return dev.mesh(shape, sp, out, res, "fd6")

    m = run()
    torch.cuda.synchronize()
    nverts, nidx = int(m[0][-1]), int(m[1][-1])
    del m
    reps = 5
    t = time.perf_counter()
    for _ in range(reps):
        m = run()
        del m
    torch.cuda.synchronize()
    gpu_ms = 1e3 * (time.perf_counter() - t) / reps
    evals = len(spans) * (res + 1) ** 3 + 7 * nverts
    # The same jobs as the app runs them: ShapeMesh::new and its first
    # update (the split test, then the 64 leaves meshed in one device batch
    # and copied into per-leaf host MeshBuffers), through the host mirror.
    from cantucci_amd.mesh import Camera, ShapeMesh
    cam = Camera((-3.0, 0.0, 0.0), (1.0, 0.0, 0.0), 1.0, 16 / 9, 0.00004, 10.0)
    ShapeMesh(shape).update(cam)
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        sm = ShapeMesh(shape)
        sm.update(cam)
        ts.append(time.perf_counter() - t)
    return {"workload": "64 leaves of [-1.2,1.2]^3 at res 64, classic(6, 2.5), whole surface nets",
            "vertices": nverts, "quads": nidx // 6, "evals": evals, "gpu_ms": gpu_ms,
            "gpu_evals_per_s": evals / gpu_ms * 1e3,
            "shape_mesh_first_update_ms": 1e3 * sorted(ts)[len(ts) // 2],
            "shape_mesh_vertices": sm.batch_timings.vertices}


def _reference_default_spans():
    """The reference application's bounding box [-1.2, 1.2]^3 pre-split twice
    (src/mesh/mod.rs:52-56): 64 leaves in create_spans order."""
    import numpy as np
    root = np.array([-1.2, -1.2, -1.2, 1.2, 1.2, 1.2], np.float32)
    spans = []
    for child in _create_spans(root):
        spans.extend(_create_spans(child))
    return np.array(spans, np.float32)


def _create_spans(span):
    """`create_spans` (src/octree/mod.rs:315-329): centre = s + (e - s) / 2;
    child bit 2 = +x half, bit 1 = +y, bit 0 = +z."""
    import numpy as np
    s, e = span[:3], span[3:]
    c = (s + (e - s) / np.float32(2.0)).astype(np.float32)
    out = []
    for i in range(8):
        lo = np.array([c[a] if (i >> (2 - a)) & 1 else s[a] for a in range(3)], np.float32)
        hi = np.array([e[a] if (i >> (2 - a)) & 1 else c[a] for a in range(3)], np.float32)
        out.append(np.concatenate([lo, hi]))
    return out


def bench_trace(ctx, args):
    """32 seeded orbit poses x 1920x1080 primary rays, 256 steps, hit d < 1e-4 t,
    escape t > 10: eager launches, and the same launch captured in a HIP graph
    and replayed.  The trace kernel's algorithmic FLOP come from the work
    counters (de_flop, 11 per step, 33 per ray)."""
    import torch

    import cantucci_amd as ca
    from cantucci_amd import autotune
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl

    autotune.apply(ctx, "trace")
    tc = wl.TRACE_1080P
    shape = ca.Mandelbulb.classic(tc["max_iters"], tc["bailout"], ctx=ctx)
    cams = torch.from_numpy(wl.orbit_poses(tc["poses"], tc["seed"], tc["W"], tc["H"], tc["fov_deg"])).cuda()
    depth = torch.empty((tc["poses"], tc["H"], tc["W"]), device="cuda")
    steps = torch.empty((tc["poses"], tc["H"], tc["W"]), dtype=torch.int32, device="cuda")

    def run():
        dev.render_depth(shape, cams, tc["W"], tc["H"], tc["max_steps"], 0.0, tc["eps_rel"], tc["t_max"],
                         depth=depth, steps=steps)

    ms = _time_ms(run, 5)
    kernel_ms = event_ms(run, 5)
    bodies, bails, evals = count_work(ctx, run)
    rays = tc["poses"] * tc["W"] * tc["H"]
    flop = de_flop(bodies, bails, evals) + FLOP_STEP * evals + FLOP_RAY * rays
    # the same launch replayed from a HIP graph on a stream of its own
    s = torch.cuda.Stream()
    ref_depth = depth.clone()
    g = dev.Graph(lambda: run(), s)
    g.launch()
    s.synchronize()
    replay_equal = bool(torch.equal(depth.nan_to_num(-1.0), ref_depth.nan_to_num(-1.0)))
    reps = 5
    t = time.perf_counter()
    for _ in range(reps):
        g.launch()
    s.synchronize()
    graph_ms = 1e3 * (time.perf_counter() - t) / reps
    g.close()
    return {"workload": "32 poses x 1920x1080, P8 15 it, 256 steps, eps 1e-4 t, t_max 10",
            "rays_per_s": rays / (ms / 1e3), "ms": ms, "ms_graph_replay": graph_ms,
            "rays_per_s_graph": rays / (graph_ms / 1e3), "graph_replay_equals_eager": replay_equal,
            "hit_fraction": float(torch.isfinite(depth).float().mean().item()),
            "mean_steps": float(steps.float().mean().item()),
            "de_evals_per_s": evals / (ms / 1e3),
            "roofline": {"kernel": "k_trace<P8Exact, RaysCamera>", "kernel_ms": kernel_ms,
                         "loop_bodies": bodies, "bailed_evals": bails, "evals": evals,
                         "algorithmic_flop_per_launch": flop, "achieved": flop / (kernel_ms / 1e3) / 1e12,
                         "unit": "TFLOP/s", "frac": flop / (kernel_ms / 1e3) / PEAK_F32}}


def bench_normals(ctx, args):
    """2^22 points in the shell 0.9 < |p| < 1.3 (seed 1), power U[6,10], 15 it:
    distance + normal by the fused forward+tangent kernel, by the reverse-sweep
    VJP, and by the 6-tap central difference (h = 1e-3)."""
    import torch

    import cantucci_amd as ca
    from cantucci_amd import autotune
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl

    sc = wl.SHELL_NORMALS
    pts, powers = wl.shell_points(sc["n"], sc["seed"])
    P = torch.from_numpy(pts).cuda()
    W = torch.from_numpy(powers).cuda()
    shape = ca.MandelbulbTrig(sc["max_iters"], sc["bailout"], 8.0, ctx=ctx)
    nrm = torch.empty((sc["n"], 3), device="cuda")
    d = torch.empty(sc["n"], device="cuda")
    autotune.apply(ctx, "tangent")
    t_tan = _time_ms(lambda: dev.normals(shape, P, W, out=nrm, d=d), 5)
    t_vjp = _time_ms(lambda: dev.de_vjp(shape, P, W), 3)
    autotune.apply(ctx, "fd")
    t_fwd = _time_ms(lambda: dev.de_points(shape, P, W), 5)
    t_fd = _time_ms(lambda: dev.normals_fd(shape, P, sc["fd_h"], W, out=nrm, d=d), 5)
    return {"workload": "2^22 shell points, power U[6,10], 15 it", "ms_forward": t_fwd,
            "ms_fwd_tangent_normals": t_tan, "ms_fwd_reverse_vjp": t_vjp, "ms_fwd_6tap_fd": t_fd,
            "speedup_tangent_vs_fd": t_fd / t_tan, "speedup_vjp_vs_fd": t_fd / t_vjp,
            "accuracy": "angular error against the FP64 dual-number oracle: "
                        "tests/test_gpu_fullsize.py::test_shell_normals_full_size"}


def bench_jobs(ctx, args):
    """64 jobs: N from {32,48,64,96,128}, power U[2,12], cap U{8..30} (seed 5);
    distances + unit normals in one persistent launch (captured in a HIP graph)."""
    import numpy as np
    import torch

    from cantucci_amd import _lib, autotune
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl

    autotune.apply(ctx, "jobs")
    N, Pw, cap = wl.mixed_jobs(**{k: v for k, v in wl.MIXED_JOBS.items() if k not in ("bailout",)})
    js = dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB_TRIG, wl.MIXED_JOBS["bailout"],
                    [dict(n=int(n), power=float(p), max_iters=int(c)) for n, p, c in zip(N, Pw, cap)])
    dists = torch.empty(js.total, device="cuda")
    nrm = torch.empty((js.total, 3), device="cuda")
    done = torch.zeros(js.njobs, dtype=torch.int64, device="cuda")
    start = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    g = dev.Graph(lambda: js.run(dists, nrm, done, start, stream=s.cuda_stream), s)
    g.launch()
    s.synchronize()
    reps = 5
    lat = []
    t = time.perf_counter()
    for _ in range(reps):
        g.launch()
        s.synchronize()
        ticks = done.cpu().numpy().astype(np.float64) - float(start.cpu().numpy()[0])
        lat.append(ticks)
    elapsed = (time.perf_counter() - t) / reps
    khz = _lib.lib.cth_wall_clock_khz(ctx.handle) or 100000
    lat_ms = np.concatenate(lat) / khz
    g.close()
    js.close()
    return {"workload": "64 mixed jobs (seed 5), trig real power, f32 distances + normals",
            "points": int(js.total), "evals_per_s": js.total / elapsed, "ms_per_batch": elapsed * 1e3,
            "job_latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                               "max": float(lat_ms.max())}}


if __name__ == "__main__":
    main()
