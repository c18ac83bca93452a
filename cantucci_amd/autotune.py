"""Launch-shape autotuner.

Every persistent lane-refill kernel takes a launch shape (cth_launch):
threads per workgroup, persistent workgroups per CU, chunk of indices per
claim, the idle-lane threshold that triggers a refill, and the schedule
(static interleave or atomic claims).  The best shape depends on the
workload's iteration-count distribution and size, so each workload is tuned
by timing a grid of shapes with HIP events and keeping the fastest; results
are cached as JSON keyed by device name and workload.

    python -m cantucci_amd.autotune [--out FILE] [--workloads spans,trace,...]
"""
from __future__ import annotations

import argparse
import itertools
import json
import time
from pathlib import Path

from . import _lib

CACHE = Path(__file__).resolve().parent / "tuned_launch.json"

SPACE = {
    "block": [256],
    "blocks_per_cu": [4, 8, 16],
    "chunk": [64, 256, 1024],
    "refill_min": [16, 32, 40, 44, 48, 56, 62, 64],
    "dynamic": [0, 1],
}

# The trace kernel batches finished evaluations' steps by a fixed threshold
# of its own; refill_min only batches the fetch of new rays, which pays off
# at small batches.
SPACES = {
    "trace": {
        "block": [256],
        "blocks_per_cu": [4, 8, 12, 16],
        "chunk": [256, 1024],
        "refill_min": [1, 2, 4, 8, 16],
        "dynamic": [0, 1],
    },
}

KERNEL_OF = {
    "grid": _lib.CTH_KERNEL_DE,
    "fd": _lib.CTH_KERNEL_DE,
    "spans": _lib.CTH_KERNEL_SPANS,
    "trace": _lib.CTH_KERNEL_TRACE,
    "tangent": _lib.CTH_KERNEL_ADJOINT,
    "jobs": _lib.CTH_KERNEL_JOBS,
    "mesh": _lib.CTH_KERNEL_MESH,
}


def device_key() -> str:
    import torch
    return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]


def time_ms(run, reps: int = 3) -> float:
    """Median of `reps` event-timed runs on torch's current stream."""
    import torch
    run()
    torch.cuda.synchronize()
    out = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        torch.cuda.synchronize()
        out.append(a.elapsed_time(b))
    out.sort()
    return out[len(out) // 2]


def configs(space: dict = SPACE):
    keys = list(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, vals))
        if cfg["dynamic"] and cfg["chunk"] < 256:
            continue
        yield cfg


def tune(ctx, kernel: int, run, space: dict = SPACE, reps: int = 3, log=None) -> tuple[dict, list]:
    """Time `run` under every shape of `space` for `kernel`; apply and return the best."""
    results = []
    for cfg in configs(space):
        ctx.set_launch(kernel, **cfg)
        ms = time_ms(run, reps)
        results.append((ms, cfg))
        if log:
            log(f"  {cfg} -> {ms:.3f} ms")
    results.sort(key=lambda t: t[0])
    best = results[0][1]
    ctx.set_launch(kernel, **best)
    return best, results


def load(path: Path = CACHE) -> dict:
    if not path.exists():
        return {}
    try:
        return json.loads(path.read_text())
    except Exception:
        return {}


def apply(ctx, workload: str, path: Path = CACHE) -> dict | None:
    """Apply the cached best shape of `workload` (for this device) to ctx."""
    entry = load(path).get(device_key(), {}).get(workload)
    if not entry:
        return None
    cfg = entry["config"]
    ctx.set_launch(KERNEL_OF[workload], **cfg)
    return cfg


def _workloads(ctx):
    """The benchmark workloads at their full sizes (each runs in milliseconds,
    so a whole sweep takes seconds)."""
    import torch

    from . import device as dev
    from . import workloads as wl
    from .shape import Mandelbulb, MandelbulbTrig

    runs = {}
    m8 = Mandelbulb.classic(15, 2.0, ctx=ctx)
    runs["grid"] = lambda: dev.de_grid(m8, (-1.5,) * 3, (1.5,) * 3, 64, 63.0)
    hd = wl.half_diagonals(3.0, 8)
    leaves, _ = dev.refine(m8, wl.OCTREE["root"], 2, 8, hd)
    out = torch.empty((leaves.shape[0], 729), device="cuda")
    runs["spans"] = lambda: dev.sample_spans(m8, leaves, 8, with_cross=False, out=out)
    dev.sample_spans(m8, leaves, 8, with_cross=False, out=out)
    runs["mesh"] = lambda: dev.mesh(m8, leaves, out, 8, "fd6")
    cams = torch.from_numpy(wl.orbit_poses(32, 0, 1920, 1080)).cuda()
    depth = torch.empty((32, 1080, 1920), device="cuda")
    steps = torch.empty((32, 1080, 1920), dtype=torch.int32, device="cuda")
    runs["trace"] = lambda: dev.render_depth(m8, cams, 1920, 1080, depth=depth, steps=steps)
    pts, powers = wl.shell_points(1 << 22, seed=1)
    P = torch.from_numpy(pts).cuda()
    W = torch.from_numpy(powers).cuda()
    tr = MandelbulbTrig(15, 2.0, 8.0, ctx=ctx)
    nrm = torch.empty((P.shape[0], 3), device="cuda")
    dd = torch.empty(P.shape[0], device="cuda")
    runs["tangent"] = lambda: dev.normals(tr, P, W, out=nrm, d=dd)
    runs["fd"] = lambda: dev.normals_fd(tr, P, 1e-3, W, out=nrm, d=dd)
    N, Pw, cap = wl.mixed_jobs(64, seed=5)
    js = dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB_TRIG, 2.0,
                    [dict(n=int(n), power=float(p), max_iters=int(c)) for n, p, c in zip(N, Pw, cap)])
    jd = torch.empty(js.total, device="cuda")
    jn = torch.empty((js.total, 3), device="cuda")
    done = torch.zeros(js.njobs, dtype=torch.int64, device="cuda")
    st = torch.zeros(1, dtype=torch.int64, device="cuda")
    runs["jobs"] = lambda: js.run(jd, jn, done, st)
    runs["_keep"] = js
    return runs


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(CACHE))
    ap.add_argument("--workloads", default="grid,spans,mesh,trace,tangent,fd,jobs")
    ap.add_argument("--mode", default="parity")
    args = ap.parse_args()
    from .shape import Context

    ctx = Context(0, args.mode)
    runs = _workloads(ctx)
    path = Path(args.out)
    cache = load(path)
    key = device_key()
    cache.setdefault(key, {})
    for name in args.workloads.split(","):
        t = time.time()
        best, results = tune(ctx, KERNEL_OF[name], runs[name], SPACES.get(name, SPACE))
        cache[key][name] = {
            "config": best,
            "best_ms": results[0][0],
            "worst_ms": results[-1][0],
            "tried": len(results),
            "mode": args.mode,
        }
        print(f"{name}: best {best} {results[0][0]:.3f} ms (worst {results[-1][0]:.3f} ms, "
              f"{len(results)} shapes, {time.time() - t:.1f} s)", flush=True)
    path.write_text(json.dumps(cache, indent=1, sort_keys=True))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
