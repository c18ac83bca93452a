"""Profiling driver: one launch of each hot kernel on its benchmark workload,
plus a plain 1 GiB device copy whose HBM counters calibrate FETCH_SIZE and
WRITE_SIZE.  Run under rocprofv3, e.g.

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o run -- python -m cantucci_amd.profile_kernels
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o fetch -- python -m cantucci_amd.profile_kernels
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT -o write -- python -m cantucci_amd.profile_kernels
"""
from __future__ import annotations

import argparse
import sys


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--legs", default="spans,mesh,trace,tangent,fd,jobs,copy")
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    legs = set(args.legs.split(","))

    import torch

    import cantucci_amd as ca
    from cantucci_amd import _lib, autotune
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl

    ctx = ca.Context(0, "parity")
    m8 = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    if "spans" in legs:
        autotune.apply(ctx, "spans")
        autotune.apply(ctx, "mesh")
        oc = wl.OCTREE
        leaves, _ = dev.refine(m8, oc["root"], oc["min_depth"], oc["max_depth"], wl.half_diagonals(3.0, 8))
        out = torch.empty((leaves.shape[0], 729), device="cuda")
        for _ in range(args.reps):
            dev.sample_spans(m8, leaves, 8, with_cross=False, out=out)
        torch.cuda.synchronize()
        print("spans done", leaves.shape[0], flush=True)
        if "mesh" in legs:
            for _ in range(args.reps):
                m = dev.mesh(m8, leaves, out, 8, "fd6")
                del m
            torch.cuda.synchronize()
            print("mesh done", flush=True)
        del out
    if "trace" in legs:
        autotune.apply(ctx, "trace")
        cams = torch.from_numpy(wl.orbit_poses(32, 0, 1920, 1080)).cuda()
        for _ in range(args.reps):
            dev.render_depth(m8, cams, 1920, 1080)
        torch.cuda.synchronize()
        print("trace done", flush=True)
    pts, powers = wl.shell_points(1 << 22, seed=1)
    P = torch.from_numpy(pts).cuda()
    W = torch.from_numpy(powers).cuda()
    tr = ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx)
    if "tangent" in legs:
        autotune.apply(ctx, "tangent")
        for _ in range(args.reps):
            dev.normals(tr, P, W)
        torch.cuda.synchronize()
        print("tangent done", flush=True)
    if "fd" in legs:
        autotune.apply(ctx, "fd")
        for _ in range(args.reps):
            dev.normals_fd(tr, P, 1e-3, W)
        torch.cuda.synchronize()
        print("fd done", flush=True)
    if "jobs" in legs:
        autotune.apply(ctx, "jobs")
        N, Pw, cap = wl.mixed_jobs(64, seed=5)
        js = dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB_TRIG, 2.0,
                        [dict(n=int(n), power=float(p), max_iters=int(c)) for n, p, c in zip(N, Pw, cap)])
        d = torch.empty(js.total, device="cuda")
        nr = torch.empty((js.total, 3), device="cuda")
        done = torch.zeros(js.njobs, dtype=torch.int64, device="cuda")
        st = torch.zeros(1, dtype=torch.int64, device="cuda")
        for _ in range(args.reps):
            js.run(d, nr, done, st)
        torch.cuda.synchronize()
        print("jobs done", flush=True)
    if "copy" in legs:
        # calibration: a plain device copy of exactly 1 GiB
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(args.reps):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copy done", flush=True)
    ctx.close()
    sys.stdout.flush()


if __name__ == "__main__":
    main()
